// Host-side context of one device.
#pragma once

#include <hip/hip_runtime_api.h>

#include <condition_variable>
#include <map>
#include <mutex>
#include <memory>
#include <tuple>
#include <vector>

#include "hddm_lb.h"
#include "layout.h"

namespace hddm_lb {

int hip_fail(hipError_t e, const char* what, const char* file, int line);
void* dev_alloc(size_t bytes, int& rc);
void dev_free(void* p);

struct DeviceGuard
{
	int prev;
	explicit DeviceGuard(int dev);
	~DeviceGuard();
};

struct StateSlot
{
	bool loaded = false;
	StateDesc desc;              // device pointers, seg fields filled per call
	std::vector<void*> bufs;     // device allocations owned by the slot
	std::vector<uint8_t> shared; // host copy: factors shared with the previous node
	std::vector<int32_t> opstart;// host copy: first op of each node
	int32_t* d_seg_op = nullptr; // device segment table (capacity nno + 1)
	int64_t seg_len = -1;        // segment length the table was built for
	// the wave-specialized adjoint's own (finer) segment table
	int32_t* d_seg_op_adj = nullptr;
	int64_t seg_len_adj = -1;
	int nseg_adj = 1;
	// every active factor has i a positive power of two (every sparse-grid
	// node): the factor table's centre -j / i is then exact, which the
	// wave-specialized adjoint's dx waves need
	bool centre_exact = true;
	~StateSlot();
	void release();
};

// Per-call device scratch of the forward sequence. The context owns one;
// every hipGraph plan owns its own, so replays never alias later eager calls.
struct Scratch
{
	uint32_t* bkey = nullptr; size_t cap_bkey = 0;
	uint32_t* perm = nullptr; size_t cap_perm = 0;
	uint32_t* hist = nullptr; size_t cap_hist = 0;
	uint32_t* cursor = nullptr; size_t cap_cursor = 0;
	double* partial = nullptr; size_t cap_partial = 0;
	double* dspart = nullptr; size_t cap_dspart = 0;   // copies of the adjoint's dS (per XCD or per owner class)
	double* dxpart = nullptr; size_t cap_dxpart = 0;   // node-segment partials of the adjoint's dx
	// the adjoint's stable point order: sorted keys, point indices, sort scratch
	uint32_t* bkey2 = nullptr; size_t cap_bkey2 = 0;
	uint32_t* pidx = nullptr; size_t cap_pidx = 0;
	char* sort_tmp = nullptr; size_t cap_sort_tmp = 0;
	StatePlan* plans = nullptr; size_t cap_plans = 0;
	PlanHeader* hdr = nullptr; size_t cap_hdr = 0;
	bool owned = true;   // false: the pointers lie in memory the scratch does not own
	~Scratch();
	void release();
};

// One host-pointer forward waiting to be coalesced with others (hddm_lb_forward_host).
struct HostRequest
{
	int64_t n;
	const double* x;
	int64_t ldx;
	const int32_t* sid;
	double* y;
	int64_t ldy;
	int prec;
	int rc;
	bool done;
};

// Launch shape of the sweep, chosen by default or by the autotuner.
struct TuneConfig
{
	int waves = 4;     // waves per workgroup
	int walk = 1;      // kWalkSmem (1) or kWalkLds (0)
	int slab = 1;      // dof tiles of 32 per sweep pass
	int ws = 0;        // 1: wave-specialized sweep (walker + consumer waves)
	int nc = 4;        // consumer waves of the wave-specialized sweep (2 or 4)
	int64_t seg = -1;  // node segment length used while tuning (-1: context option)
};

} // namespace hddm_lb

struct hddm_lb_ctx
{
	int device = 0;
	int dim = -1;
	int num_cus = 0;
	size_t max_lds = 0;
	size_t total_mem = 0;        // device memory, bytes
	std::vector<std::unique_ptr<hddm_lb::StateSlot> > states;

	// options
	bool opt_autotune = false;
	int64_t opt_seg_nodes = 0;
	int opt_waves = 0;
	bool opt_sort = true;
	bool opt_timing = false;
	int opt_slab = 0;
	int opt_walk = 0;      // 0 automatic, 1 scalar loads, 2 LDS tiles
	int opt_form = 0;      // 0 automatic, 1 one wave walks and contracts, 2 wave-specialized

	// device tables
	bool tables_dirty = true;
	int tables_n = 0, tables_cap = 0;
	int dt_needed = 1, fmax_all = 1;
	bool centre_exact_all = true;   // every state's StateSlot::centre_exact
	int64_t max_nseg_dp = 0;
	int max_nseg = 1;            // most node segments of a split state (1: none is split)
	hddm_lb::StateDesc* d_states = nullptr;
	// the same descriptors with the wave-specialized adjoint's segment tables
	hddm_lb::StateDesc* d_states_adj = nullptr;
	int max_nseg_adj = 1;
	int64_t nseg_adj_total = 0;  // segments of all loaded states in that table
	int32_t* d_cost_order = nullptr;
	uint8_t* d_state_ok = nullptr;
	uint64_t generation = 0;     // bumped whenever states or tables change

	// per-call scratch of eager forwards
	hddm_lb::Scratch scratch;
	double* d_xstage = nullptr; size_t cap_xstage = 0;
	double* d_ystage = nullptr; size_t cap_ystage = 0;
	int32_t* d_sidstage = nullptr; size_t cap_sidstage = 0;
	hipStream_t own_stream = nullptr;
	// host-pointer calls: pinned host staging (hipHostMalloc) and, for small
	// batches, captured forwards keyed by (points, ldx, ldy, state ids given,
	// precision) that replay with one graph launch
	double* h_xstage = nullptr; size_t cap_hx = 0;
	double* h_ystage = nullptr; size_t cap_hy = 0;
	int32_t* h_sidstage = nullptr; size_t cap_hsid = 0;
	std::map<std::tuple<int64_t, int64_t, int64_t, int, int>, hddm_lb_plan*> small_plans;
	// host-pointer calls from several threads: run_mu serializes the staging
	// and launch; small calls queue in comb_queue and one of the waiting
	// threads runs them as one batched forward
	std::mutex run_mu, comb_mu;
	std::condition_variable comb_cv;
	std::vector<hddm_lb::HostRequest*> comb_queue;
	bool comb_busy = false;
	bool opt_coalesce = true;
	bool opt_device_layout = true;   // build uploaded layouts on the device
	bool opt_adj_owned = false;      // wave-specialized adjoint: dS through owned units (bitwise repeatable)
	int64_t comb_requests = 0, comb_batches = 0;
	// Eager calls share `scratch`: each one records scratch_event on its
	// stream, and the next call on another stream waits for it first.
	hipEvent_t scratch_event = nullptr;
	hipStream_t scratch_stream = nullptr;
	bool scratch_pending = false;

	// adjoint scratch
	float* d_wbuf = nullptr; size_t cap_wbuf = 0;

	// measurement
	bool opt_stats = false;
	// walk ops, prunes, chunks contracted, work items, MFMA instructions issued, nnz pairs
	static constexpr int kStatCount = 6;
	unsigned long long* d_stats = nullptr;
	uint32_t* d_status = nullptr;   // HDDM_LB_STATUS_* bits raised by kernels
	// device address of this device's copy of the allocation-free entry
	// point's static scratch array (hipGetSymbolAddress at create)
	char* noalloc_scratch = nullptr;
	// wave-specialized adjoint in statistics mode: walker cycles in walk A,
	// walk B, waiting for w, waiting for ring slots; consumer cycles in w, in
	// dS, idle; items; chunks
	static constexpr int kAdjStatCount = 9;
	unsigned long long* d_adj_stats = nullptr;
	// rings of event pairs bracketing the kernels of timed calls: [0] the
	// sweep kernels of forwards, [1] the adjoint kernels
	static constexpr int kTimingRing = 256;
	struct TimingRing
	{
		std::vector<hipEvent_t> e0, e1;
		int64_t calls = 0;   // pairs recorded since create
		int64_t read = 0;    // pairs already summarised
	};
	TimingRing tring[2];
	bool last_timed = false;

	// autotuner cache: (prec, dim, total nodes, max dofs, log2 npoints) -> config
	// (precision, D, total nodes, max dofs, log2 points, the options that
	// restrict the candidates: formulation, waves, slab tiles)
	std::map<std::tuple<int, int, int64_t, int, int, int, int, int>, hddm_lb::TuneConfig> tuned;
	int64_t tune_runs = 0, tune_hits = 0;   // keys tuned, forwards served from the cache
	hddm_lb::TuneConfig last_cfg;           // configuration of the last eager forward
	int last_adj[4] = { 0, 0, 0, 0 };       // hddm_lb_last_adjoint_config

	~hddm_lb_ctx();
};

namespace hddm_lb {

int prepare_tables(hddm_lb_ctx* ctx);
// Looks up this device's copy of the allocation-free scratch array (noalloc.hip).
int noalloc_init(hddm_lb_ctx* ctx);
// Orders a call on st after the last eager call that used ctx->scratch on another stream.
int scratch_wait(hddm_lb_ctx* ctx, hipStream_t st);
// Marks the end of an eager call's use of ctx->scratch on st.
int scratch_mark(hddm_lb_ctx* ctx, hipStream_t st);
int prepare_scratch(hddm_lb_ctx* ctx, Scratch& sc, int64_t npoints);
TuneConfig default_config(hddm_lb_ctx* ctx, int prec);
int enqueue_forward(hddm_lb_ctx* ctx, Scratch& sc, int64_t npoints, const double* x, int64_t ldx,
	const int32_t* sid, double* y, int64_t ldy, int prec, hipStream_t st, const TuneConfig& cfg, bool timed);
int autotune(hddm_lb_ctx* ctx, int64_t npoints, const double* x, int64_t ldx, const int32_t* sid,
	double* y, int64_t ldy, int prec, hipStream_t st, TuneConfig& cfg);
int bucket_bits(int64_t npoints, int nstates, int sort);
int check_forward_args(hddm_lb_ctx* ctx, int64_t npoints, const double* x, int64_t ldx,
	double* y, int64_t ldy, int prec);

} // namespace hddm_lb
