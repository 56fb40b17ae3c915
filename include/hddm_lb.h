/*
 * hddm_lb.h -- flat C ABI of the MI355X (gfx950) LinearBasis interpolation engine.
 *
 * Every entry point takes plain pointers and sizes; no C++ or torch types cross
 * this boundary. Device-side entry points take device pointers and a
 * hipStream_t passed as void*, are stream-ordered and return a status code;
 * nothing aborts the process. hddm_lb_last_error() returns the text of the
 * last failure on the calling thread.
 *
 * The reference interfaces these entry points replace (paths relative to the
 * reference checkout; LB = LinearBasis):
 *   - include/interpolator.h:31-33   Interpolator::interpolate(device, data, istate, x, DofPerNode, value)
 *       -> hddm_lb_forward / hddm_lb_forward_host (one point = npoints 1)
 *   - include/interpolator.h:35-37   Interpolator::interpolate(device, data, x[], DofPerNode, value[])
 *       -> hddm_lb_forward with a state id per point
 *   - LB/gold/src/InterpolateArray.cpp:8-43 and LB/cuda/src/InterpolateArray.cu:239-267
 *       (the extern "C" kernels the Interpolator dispatches to) -> hddm_lb_forward
 *   - LB/gold/src/Data.cpp:322-595 DataDense::load (.plt reader)
 *       -> hddm_lb_plt_read + hddm_lb_upload_state
 *   - LB/cuda/src/Data.cu:944-948 (setXPS/setChains/setSurplus device upload)
 *       -> hddm_lb_upload_state
 *   - LB/cuda/src/Devices.cu:13-47 (GPU count) -> hddm_lb_device_count
 * The adjoint, the batched point API, the hipGraph plan and the grid generator
 * have no counterpart in the reference.
 */
#ifndef HDDM_LB_H
#define HDDM_LB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hddm_lb_ctx hddm_lb_ctx;
typedef struct hddm_lb_plan hddm_lb_plan;
typedef struct hddm_lb_grid hddm_lb_grid;

/* status codes */
#define HDDM_LB_SUCCESS               0
#define HDDM_LB_ERROR_INVALID_VALUE   1
#define HDDM_LB_ERROR_NO_DEVICE       2
#define HDDM_LB_ERROR_OUT_OF_MEMORY   3
#define HDDM_LB_ERROR_HIP             4
#define HDDM_LB_ERROR_STATE           5
#define HDDM_LB_ERROR_IO              6
#define HDDM_LB_ERROR_FORMAT          7
#define HDDM_LB_ERROR_UNSUPPORTED     8

/* arithmetic of the contraction (the basis is always formed in FP64) */
#define HDDM_LB_FP64    0  /* FP64 VALU contraction, FP64 accumulation */
#define HDDM_LB_F16X2   1  /* MFMA f16: basis hi+lo x scaled surplus hi+lo, 3 products */
#define HDDM_LB_BF16X3  2  /* MFMA bf16: 3-term splits, 6 products */
#define HDDM_LB_F32     3  /* MFMA f32: basis f32 x surplus f32 */

/* options for hddm_lb_set_option */
#define HDDM_LB_OPT_AUTOTUNE      1  /* 1: on first use of a (precision, D, nodes, dofs, log2 points) key, time
                                        the sweep formulations and launch shapes and keep the fastest */
#define HDDM_LB_OPT_SEGMENT_NODES 2  /* node-range length of one work item (0: automatic, eighths of states
                                        of >= 16k nodes; a value >= nno keeps states whole) */
#define HDDM_LB_OPT_WAVES_PER_WG  3  /* waves per workgroup (1, 2, 4; 0: automatic) */
#define HDDM_LB_OPT_SORT_POINTS   4  /* 1 (default): spatially sort points before the sweep */
#define HDDM_LB_OPT_TIMING        5  /* 1: time the sweep kernels with hipEvents (hddm_lb_last_sweep_ms) */
#define HDDM_LB_OPT_SLAB_TILES    6  /* dof tiles of 32 per sweep pass (1, 2; 0: automatic) */
/* option 7 is unused: MFMA partial sums are folded into FP64 after every chunk */
#define HDDM_LB_OPT_STATS         8  /* 1: count walk ops, prunes, contracted chunks and items (hddm_lb_last_stats) */
#define HDDM_LB_OPT_WALK          9  /* op stream source: 0 automatic, 1 scalar loads per wave, 2 LDS tiles per workgroup */
#define HDDM_LB_OPT_FORMULATION  10  /* forward and adjoint: 0 automatic, 1 each wave walks and contracts
                                        its block (per-tile kernels), 2 a walker wave feeds MFMA consumer
                                        waves (wave-specialized kernels). The wave-specialized adjoint
                                        takes states whose index_i values are powers of two (every sparse
                                        grid) and, when dim > 32, nodes of at most 8 factors; 0 runs
                                        other states on the per-tile kernels, 2 fails for them */
#define HDDM_LB_OPT_COALESCE     11  /* 1 (default): hddm_lb_forward_host calls of at most 64 points made
                                        at the same time from several threads run as one batched forward */
#define HDDM_LB_OPT_DEVICE_LAYOUT 12 /* 1 (default): hddm_lb_upload_state builds the node layout (trie
                                        order, op stream, surplus splits, factor table) on the device,
                                        bit for bit as the host build; 0: on the host */
#define HDDM_LB_OPT_ADJOINT_OWNED 13 /* 1: the wave-specialized adjoint adds dS through owned units
                                        (one workgroup per row of one of K dS copies, plain loads and
                                        stores, the copies summed in a fixed order): dS bitwise
                                        repeatable, when the copies fit (at most 16 GiB and an eighth
                                        of device memory); slower than atomics at D=59 (one 65,536-
                                        point F32 call 220.5 against 175.4 ms). 0 (default): FP64 atomics */

/* bits of the device status word (hddm_lb_device_status) */
#define HDDM_LB_STATUS_WAIT_LIMIT 1u /* a wave of a wave-specialized kernel stopped waiting for another
                                        wave of its workgroup: the results of that call are invalid */
#define HDDM_LB_STATUS_WALK_MISMATCH 2u /* the adjoint's two walks of one trie disagreed on the surviving
                                           nodes: the dx of that call is invalid */

/* ---- devices and contexts ---- */

/* Number of visible HIP devices; 0 (and success) when there is none. */
int hddm_lb_device_count(int* count);

/* One context per device per host thread; allocates nothing until states are uploaded. */
int hddm_lb_create(int device, hddm_lb_ctx** ctx);
int hddm_lb_destroy(hddm_lb_ctx* ctx);

/* Text of the last error on this thread (never NULL). */
const char* hddm_lb_last_error(void);

/* Waits for the device and returns the status bits kernels of this context
 * have raised since the last clearing call (0: none); clear = 1 resets them. */
int hddm_lb_device_status(hddm_lb_ctx* ctx, uint32_t* status, int clear);

int hddm_lb_set_option(hddm_lb_ctx* ctx, int option, int64_t value);

/* Upload one state from the reference's dense representation: for node k and
 * dim d, index_i[k*ld_index + d] is i (0 for level 1, 2^(l-1) otherwise) and
 * index_j[k*ld_index + d] is j, exactly as the gold Matrix<int> holds them
 * (LB/gold/src/Data.cpp:474-494); surplus is row-major [nno x ld_surplus]
 * FP64. Host pointers. All states of a context share dim (the reference's
 * rule, LB/gold/src/Data.cpp:424-429); dofs may differ per state. Replacing
 * an uploaded state is allowed. */
int hddm_lb_upload_state(hddm_lb_ctx* ctx, int istate, int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus);

/* Copy of one device array of an uploaded state's layout (for checks): 0 op
 * stream (with its sentinels), 1 skip, 2 trie order, 3 FP64 surplus rows,
 * 4 f16 hi/lo chunks, 5 bf16 chunks, 6 f32 chunks, 7 FP32 surplus rows,
 * 8 f16 column scales, 9 and 10 support masks, 11 factor table; in the layouts
 * hddm_lb_layout_* describe. With dst NULL, *bytes receives the size; else
 * *bytes (at least the size) bytes of dst are available and receive it. */
int hddm_lb_state_array(hddm_lb_ctx* ctx, int istate, int array, void* dst, uint64_t* bytes);
int hddm_lb_state_info(hddm_lb_ctx* ctx, int istate,
	int* dim, int* nno, int* dofs, int* max_factors);

/* Forward interpolation of npoints points (device pointers):
 *   x         [npoints x ldx] FP64 coordinates in [0, 1]
 *   state_ids [npoints] int32 state of each point, or NULL for state 0
 *   y         [npoints x ldy] FP64; row p receives dofs(state_ids[p]) values
 * Stream-ordered on `stream` (a hipStream_t, NULL for the default stream). */
int hddm_lb_forward(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream);

/* hddm_lb_forward without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc). Same arguments,
 * results and stream semantics as hddm_lb_forward with the default sweep
 * configuration (no autotuning). All per-call scratch comes from one static
 * device array of 64 MiB per device, shared by every context on that device:
 * calls that overlap on one device race. A batch whose scratch (8 bytes per
 * point, plus 8 * nseg * dofs_pad bytes per point for states split into node
 * segments) exceeds 64 MiB is rejected with HDDM_LB_ERROR_INVALID_VALUE;
 * statistics mode is rejected; the state tables must be ready (they are after
 * every upload and option change). */
int hddm_lb_forward_noalloc(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream);

/* Same contract with host pointers; synchronous. This is what the legacy
 * per-point Interpolator::interpolate maps onto. Thread-safe on one context:
 * x and y go through pinned staging buffers (one DMA each way); a call of at
 * most 64 points replays a captured forward (one graph launch), and calls of
 * at most 64 points made concurrently from several threads are queued and
 * run as one batched forward (HDDM_LB_OPT_COALESCE). */
int hddm_lb_forward_host(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision);

/* Adjoint (device pointers; no counterpart in the reference). For the
 * forward y_p = sum_k B_k(x_p) S_k and gbar [npoints x ldg] FP64 = dL/dy:
 *   dx        [npoints x lddx] FP64: dx_p = sum_k (gbar_p . S_k) dB_k/dx_p with the
 *             piecewise-constant hat derivative -i sign(x i - j) on the support
 *   dsurplus  FP64, state s at element offset hddm_lb_surplus_offset(s), rows in
 *             the uploaded node order, row length dofs(s): dS_k = sum_p B_k(x_p) gbar_p
 * Both outputs are overwritten (dx only in its first dim columns, and only in
 * rows whose state id is valid); either may be NULL. The basis is recomputed.
 * dx is summed in a fixed order (bitwise repeatable). dsurplus is bitwise
 * repeatable on the wave-specialized kernel with owned units
 * (HDDM_LB_OPT_ADJOINT_OWNED); elsewhere it is summed with FP64 atomics and
 * its last bits may vary between runs.
 * precision: HDDM_LB_FP64, or HDDM_LB_F32 (held to 1e-6 normwise: on the
 * wave-specialized kernel the products g.S and B^T g run on the bf16 MFMA
 * units with each F32 operand split into three bf16 parts (F32 accuracy),
 * F32 accumulation per 64-point chunk and FP64 accumulation beyond; the
 * per-tile kernels compute in FP64 for either precision). Stream-ordered. */
int hddm_lb_adjoint(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	const double* gbar, int64_t ldg, double* dx, int64_t lddx,
	double* dsurplus, int precision, void* stream);

int hddm_lb_surplus_offset(hddm_lb_ctx* ctx, int istate, int64_t* offset, int64_t* total);

/* ---- hipGraph plans: a forward with fixed pointers and sizes, captured once
 * and replayed; every launch of the forward sequence is in the graph. ---- */
int hddm_lb_plan_create(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream, hddm_lb_plan** plan);
int hddm_lb_plan_launch(hddm_lb_plan* plan, void* stream);
int hddm_lb_plan_destroy(hddm_lb_plan* plan);

/* ---- measurement ---- */
/* Time in ms between hipEvents recorded on the forward's stream right before
 * the first and right after the last sweep kernel of the last hddm_lb_forward
 * on this context, when HDDM_LB_OPT_TIMING is set; -1 otherwise. Waits for
 * the second event. */
double hddm_lb_last_sweep_ms(hddm_lb_ctx* ctx);
/* Sweep-kernel times of the timed forwards enqueued since the previous call
 * (at most the last 256): count, total, min and max in ms (min/max may be
 * NULL). Waits for their events. */
int hddm_lb_sweep_time_summary(hddm_lb_ctx* ctx, int64_t* ncalls, double* total_ms,
	double* min_ms, double* max_ms);
/* The same for one kind of kernel: 0 the sweep kernels of forwards (as
 * above), 1 the adjoint kernels of hddm_lb_adjoint (from before the first to
 * after the last adjoint kernel, the final reductions excluded). */
int hddm_lb_kernel_time_summary(hddm_lb_ctx* ctx, int kind, int64_t* ncalls, double* total_ms,
	double* min_ms, double* max_ms);
/* With HDDM_LB_OPT_STATS set: counts of the last forward, summed over all
 * work items: stats[0] ops evaluated, [1] subtrees pruned, [2] chunks
 * contracted (8 nodes on the FP64 path, 32 on F32, 16 on F16X2 / BF16X3),
 * [3] work items. Waits for the forward to finish. */
int hddm_lb_last_stats(hddm_lb_ctx* ctx, uint64_t* stats);
/* The first n (1 to 6) of the same counters; stats[4] is the number of MFMA
 * instructions the contraction issued (v_mfma_f32_32x32x2f32 on F32,
 * v_mfma_f32_32x32x16_{f16,bf16} on the split paths, v_mfma_f64_16x16x4f64
 * on the FP64 path of the wave-specialized sweep, 0 on the FP64 VALU path);
 * stats[5] is the number of (point, node) pairs with a non-zero basis value
 * that the contraction received (wave-specialized sweep only, else 0). */
int hddm_lb_last_stats_n(hddm_lb_ctx* ctx, uint64_t* stats, int n);

/* With HDDM_LB_OPT_STATS set, counters of the last wave-specialized adjoint,
 * summed over workgroups (shader-clock cycles from s_memtime): [0] walker in
 * the walk, [1] dx wave adding dx terms, [2] dx wave waiting for messages and
 * g.S, [3] walker waiting for ring slots, [4] consumers computing g.S,
 * [5] consumers computing dS, [6] consumers idle, [7] work items, [8] chunks.
 * The first n (1 to 9). Waits for the device. */
int hddm_lb_last_adjoint_stats(hddm_lb_ctx* ctx, uint64_t* stats, int n);

/* Host-pointer calls served by coalescing (requests) and the batched
 * forwards that served them (batches), since the context was created. */
int hddm_lb_host_call_stats(hddm_lb_ctx* ctx, int64_t* requests, int64_t* batches);

/* Sweep configuration of the last hddm_lb_forward: config[0] formulation
 * (1 each wave walks and contracts, 2 wave-specialized), [1] waves per
 * workgroup, [2] dof tiles of 32 per launch, [3] keys the autotuner has
 * tuned on this context, [4] forwards it served from its cache. */
int hddm_lb_last_config(hddm_lb_ctx* ctx, int* config);

/* Configuration of the last hddm_lb_adjoint: config[0] kernels (1 per-tile,
 * 2 wave-specialized), [1] dS copies of owned units (0: dS added with
 * atomics), [2] dS copies summed after the kernels (1: none), [3] most node
 * segments of a state in the kernels' tables. */
int hddm_lb_last_adjoint_config(hddm_lb_ctx* ctx, int* config);

/* ---- host-side node layout, for inspection without a device ---- */
typedef struct hddm_lb_layout hddm_lb_layout;
/* Builds the layout hddm_lb_upload_state would upload: nodes in prefix-trie
 * order (lexicographic by their ascending-dim factor sequences) and the op
 * stream of the walk, one op per factor a node does not share with its
 * predecessor. Op q is 4 words: 2 * dim | (skip[q] - q) << 8; ctrl = depth
 * | (last factor of its node) << 4 | (relation to op q-1: 0 first child,
 * 1 sibling, 2 other) << 5 | (op has children) << 7 | node << 8; and the high
 * words of the FP64 values of i and -j (low words 0). skip[q] is the first op
 * after the run of nodes sharing op q's prefix; opstart[n] (n <= nno) is node
 * n's first op. */
int hddm_lb_layout_build(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, hddm_lb_layout** layout);
int hddm_lb_layout_info(const hddm_lb_layout* layout, int* max_factors, int* nops,
	int* nchunks, int* dofs_pad);
/* ops: 4 words per op as above (4 * nops); skip: nops; opstart: nno + 1 */
int hddm_lb_layout_ops(const hddm_lb_layout* layout, uint32_t* ops, uint32_t* skip, int32_t* opstart);
/* order[n]: uploaded index of the node at trie position n */
int hddm_lb_layout_order(const hddm_lb_layout* layout, int32_t* order);
/* The per-node factor table the adjoint's dx terms read: fstride (4, 8 or 16)
 * slots of 4 words per node in trie order, nodes rounded up to 32; a slot is
 * {2 * dim, high word of i, high word of -j, 1}, padding slots are zero.
 * ftab may be NULL (stride only). */
int hddm_lb_layout_factors(const hddm_lb_layout* layout, uint32_t* ftab, int* fstride);
int hddm_lb_layout_free(hddm_lb_layout* layout);

/* ---- .plt surplus files (LB/gold/src/Data.cpp:13-595) ---- */
/* Reads a compressed or uncompressed (binary or text) .plt file into newly
 * allocated dense arrays (free with hddm_lb_free). nagents plays the role of
 * the config's nagents: an uncompressed file whose first int differs from it
 * is read as text, and a dim mismatch is an error, as in the reference. */
int hddm_lb_plt_read(const char* path, int nagents, int* dim, int* nno, int* dofs,
	int* level, int32_t** index_i, int32_t** index_j, double** surplus);
/* Compressed writer: the layout the reference reads at LB/gold/src/Data.cpp:50-190,525-572. */
int hddm_lb_plt_write_compressed(const char* path, int dim, int nno, int dofs, int level,
	const int32_t* index_i, const int32_t* index_j, const double* surplus);
/* Uncompressed writer from levels and indices (binary = 1, text = 0). */
int hddm_lb_plt_write_uncompressed(const char* path, int binary, int dim, int nno, int dofs,
	int level, const int32_t* levels, const int32_t* indices, const double* surplus);
void hddm_lb_free(void* p);

/* ---- seeded sparse grids (synthetic inputs) ---- */
/* Regular grid: all nodes with sum_d (l_d - 1) <= start_sum, then rounds of
 * random refinement (SplitMix64, seed): every unrefined node with
 * sum_d (l_d - 1) < max_level - 1 is refined with probability p_refine
 * (all its hierarchical children in every dim are added) until the grid holds
 * at least target_nno nodes (0: no target) or nothing is left to refine. */
int hddm_lb_grid_generate(int dim, int start_sum, int max_level, double p_refine,
	uint64_t seed, int64_t target_nno, hddm_lb_grid** grid);
int hddm_lb_grid_nno(const hddm_lb_grid* grid, int64_t* nno);
/* Dense (i, j) as uploaded, and (level, index) as stored in uncompressed .plt files. */
int hddm_lb_grid_index(const hddm_lb_grid* grid, int32_t* index_i, int32_t* index_j);
int hddm_lb_grid_levels(const hddm_lb_grid* grid, int32_t* levels, int32_t* indices);
int hddm_lb_grid_free(hddm_lb_grid* grid);

#ifdef __cplusplus
}
#endif

#endif /* HDDM_LB_H */
