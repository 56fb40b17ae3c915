// C ABI: contexts, state upload, forward launch sequencing, hipGraph plans.
#include "hddm_lb.h"
#include "errors.h"
#include "kernels.h"
#include "layout_build.h"
#include "context.h"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

namespace hddm_lb {

char* error_buffer()
{
	static thread_local char buf[1024] = "no error";
	return buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line)
{
	return set_error(HDDM_LB_ERROR_HIP, "%s: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

void* dev_alloc(size_t bytes, int& rc)
{
	void* p = nullptr;
	hipError_t e = hipMalloc(&p, bytes ? bytes : 16);
	if (e != hipSuccess)
	{
		(void)hipGetLastError();
		rc = set_error(HDDM_LB_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
		return nullptr;
	}
	rc = HDDM_LB_SUCCESS;
	return p;
}

void dev_free(void* p)
{
	if (p) (void)hipFree(p);
}

StateSlot::~StateSlot() { release(); }

Scratch::~Scratch() { release(); }

void Scratch::release()
{
	if (owned)
		for (void* p : { (void*)bkey, (void*)perm, (void*)hist, (void*)cursor, (void*)partial, (void*)dspart,
				(void*)dxpart, (void*)plans, (void*)hdr, (void*)bkey2, (void*)pidx, (void*)sort_tmp })
			dev_free(p);
	bkey = perm = hist = cursor = nullptr;
	partial = dspart = dxpart = nullptr;
	bkey2 = pidx = nullptr;
	sort_tmp = nullptr;
	cap_dspart = cap_dxpart = 0;
	cap_bkey2 = cap_pidx = cap_sort_tmp = 0;
	plans = nullptr;
	hdr = nullptr;
	cap_bkey = cap_perm = cap_hist = cap_cursor = cap_partial = cap_plans = cap_hdr = 0;
}

void StateSlot::release()
{
	for (void* p : bufs) dev_free(p);
	bufs.clear();
	loaded = false;
}

DeviceGuard::DeviceGuard(int dev) : prev(-1)
{
	if (hipGetDevice(&prev) != hipSuccess) prev = -1;
	if (prev != dev) (void)hipSetDevice(dev);
}

DeviceGuard::~DeviceGuard()
{
	int cur = -1;
	if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
}

} // namespace hddm_lb

using namespace hddm_lb;

#define HIP_TRY(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return hip_fail(e_, #expr, __FILE__, __LINE__); } while (0)

hddm_lb_ctx::~hddm_lb_ctx()
{
	DeviceGuard g(device);
	for (auto& kv : small_plans) hddm_lb_plan_destroy(kv.second);
	small_plans.clear();
	for (void* p : { (void*)h_xstage, (void*)h_ystage, (void*)h_sidstage })
		if (p) (void)hipHostFree(p);
	states.clear();
	scratch.release();
	for (void* p : { (void*)d_states, (void*)d_states_adj, (void*)d_cost_order, (void*)d_state_ok,
		(void*)d_xstage, (void*)d_ystage, (void*)d_sidstage, (void*)d_wbuf, (void*)d_stats, (void*)d_status, (void*)d_adj_stats })
		dev_free(p);
	if (scratch_event) (void)hipEventDestroy(scratch_event);
	for (auto& r : tring)
	{
		for (hipEvent_t e : r.e0) (void)hipEventDestroy(e);
		for (hipEvent_t e : r.e1) (void)hipEventDestroy(e);
	}
	if (own_stream) (void)hipStreamDestroy(own_stream);
}

extern "C" const char* hddm_lb_last_error(void) { return error_buffer(); }

extern "C" int hddm_lb_device_status(hddm_lb_ctx* ctx, uint32_t* status, int clear)
{
	if (!ctx || !status) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "device_status: NULL argument");
	DeviceGuard g(ctx->device);
	HIP_TRY(hipDeviceSynchronize());
	HIP_TRY(hipMemcpy(status, ctx->d_status, sizeof(uint32_t), hipMemcpyDeviceToHost));
	if (clear) HIP_TRY(hipMemset(ctx->d_status, 0, sizeof(uint32_t)));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_device_count(int* count)
{
	if (!count) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "device_count: count is NULL");
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) { (void)hipGetLastError(); n = 0; }
	*count = n;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_create(int device, hddm_lb_ctx** out)
{
	if (!out) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "create: ctx is NULL");
	*out = nullptr;
	int n = 0;
	hddm_lb_device_count(&n);
	if (n == 0) return set_error(HDDM_LB_ERROR_NO_DEVICE, "create: no HIP device is visible");
	if (device < 0 || device >= n)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "create: device %d out of range [0, %d)", device, n);
	DeviceGuard g(device);
	std::unique_ptr<hddm_lb_ctx> ctx(new hddm_lb_ctx);
	ctx->device = device;
	int cus = 0, lds = 0;
	HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
	HIP_TRY(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device));
	ctx->num_cus = cus;
	ctx->max_lds = (size_t)lds;
	{
		size_t mem = 0;
		HIP_TRY(hipDeviceTotalMem(&mem, device));
		ctx->total_mem = mem;
	}
	HIP_TRY(hipEventCreateWithFlags(&ctx->scratch_event, hipEventDisableTiming));
	{
		int rc = 0;
		ctx->d_status = (uint32_t*)dev_alloc(sizeof(uint32_t), rc);
		if (rc) return rc;
		HIP_TRY(hipMemset(ctx->d_status, 0, sizeof(uint32_t)));
	}
	{
		int rc = noalloc_init(ctx.get());
		if (rc) return rc;
	}
	for (auto& r : ctx->tring)
		for (int i = 0; i < hddm_lb_ctx::kTimingRing; i++)
		{
			hipEvent_t e0 = nullptr, e1 = nullptr;
			HIP_TRY(hipEventCreate(&e0));
			r.e0.push_back(e0);
			HIP_TRY(hipEventCreate(&e1));
			r.e1.push_back(e1);
		}
	*out = ctx.release();
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_destroy(hddm_lb_ctx* ctx)
{
	delete ctx;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_set_option(hddm_lb_ctx* ctx, int option, int64_t value)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: ctx is NULL");
	switch (option)
	{
	case HDDM_LB_OPT_AUTOTUNE: ctx->opt_autotune = value != 0; break;
	case HDDM_LB_OPT_SEGMENT_NODES:
		if (value < 0) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: segment nodes must be >= 0");
		ctx->opt_seg_nodes = value; ctx->tables_dirty = true; ctx->generation++;
		if (ctx->dim > 0) return prepare_tables(ctx);
		break;
	case HDDM_LB_OPT_WAVES_PER_WG:
		if (!(value == 0 || value == 1 || value == 2 || value == 4))
			return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: waves per workgroup must be 0, 1, 2 or 4");
		ctx->opt_waves = (int)value; break;
	case HDDM_LB_OPT_SORT_POINTS: ctx->opt_sort = value != 0; break;
	case HDDM_LB_OPT_TIMING: ctx->opt_timing = value != 0; break;
	case HDDM_LB_OPT_SLAB_TILES:
		if (!(value == 0 || value == 1 || value == 2))
			return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: slab tiles must be 0, 1 or 2");
		ctx->opt_slab = (int)value; break;
	case HDDM_LB_OPT_STATS: ctx->opt_stats = value != 0; break;
	case HDDM_LB_OPT_WALK:
		if (value < 0 || value > 2) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: walk must be 0, 1 or 2");
		ctx->opt_walk = (int)value; break;
	case HDDM_LB_OPT_COALESCE: ctx->opt_coalesce = value != 0; break;
	case HDDM_LB_OPT_DEVICE_LAYOUT: ctx->opt_device_layout = value != 0; break;
	case HDDM_LB_OPT_ADJOINT_OWNED:
		if (ctx->opt_adj_owned != (value != 0))
		{
			// the adjoint's segment tables depend on it
			ctx->opt_adj_owned = value != 0;
			ctx->tables_dirty = true;
			ctx->generation++;
			if (ctx->dim > 0) return prepare_tables(ctx);
		}
		break;
	case HDDM_LB_OPT_FORMULATION:
		if (value < 0 || value > 2) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: formulation must be 0, 1 or 2");
		ctx->opt_form = (int)value; break;
	default:
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "set_option: unknown option %d", option);
	}
	return HDDM_LB_SUCCESS;
}

static int finish_upload(hddm_lb_ctx* ctx, int istate, int dim, std::unique_ptr<StateSlot> slot);

extern "C" int hddm_lb_upload_state(hddm_lb_ctx* ctx, int istate, int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "upload_state: ctx is NULL");
	if (istate < 0 || istate > (1 << 14))
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "upload_state: state %d out of range [0, 16384]", istate);
	int loaded_others = 0;
	for (size_t s = 0; s < ctx->states.size(); s++)
		if ((int)s != istate && ctx->states[s] && ctx->states[s]->loaded) loaded_others++;
	if (loaded_others && dim != ctx->dim)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE,
			"upload_state: # of dimensions (%d) mismatches another state (%d)", dim, ctx->dim);
	if (xbanks_for(dim) == 0)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "upload_state: dim %d exceeds 64 (x is held in registers)", dim);

	DeviceGuard g(ctx->device);
	std::unique_ptr<StateSlot> slot(new StateSlot);
	int rc = HDDM_LB_SUCCESS;
	StateDesc& d = slot->desc;
	if (index_i && index_j && ld_index >= dim)
		for (int64_t k = 0; k < nno && slot->centre_exact; k++)
			for (int t = 0; t < dim; t++)
			{
				const int32_t i = index_i[k * ld_index + t], j = index_j[k * ld_index + t];
				if ((i || j) && !(i > 0 && (i & (i - 1)) == 0)) { slot->centre_exact = false; break; }
			}
	std::memset(&d, 0, sizeof(d));
	d.dim = dim; d.nno = nno; d.dofs = dofs;
	d.nseg = 1;
	// The layout is built on the device unless HDDM_LB_OPT_DEVICE_LAYOUT is 0;
	// inputs the device build does not take go to the host build, which builds
	// them or reports the error.
	if (ctx->opt_device_layout)
	{
		DeviceLayout DL;
		auto alloc = [&](size_t bytes, int& arc) -> void* {
			if (arc) return nullptr;
			void* p = dev_alloc(bytes, arc);
			if (p) slot->bufs.push_back(p);
			return p;
		};
		rc = build_layout_device(dim, nno, dofs, index_i, index_j, ld_index, surplus, ld_surplus, alloc, DL);
		if (rc != kLayoutHostCheck)
		{
			if (rc) return rc;
			d.dofs_pad = DL.dofs_pad; d.nops = DL.nops; d.nchunks = DL.nchunks; d.max_depth = DL.fmax;
			d.ops = DL.ops; d.skip = DL.skip; d.s64 = DL.s64; d.s16 = DL.s16; d.sbf = DL.sbf; d.s32 = DL.s32;
			d.s32r = DL.s32r; d.ftab = DL.ftab; d.fstride = DL.fstride; d.iscale = DL.iscale; d.order = DL.order;
			d.lomask = DL.lomask; d.himask = DL.himask;
			slot->shared = std::move(DL.shared);
			slot->opstart = std::move(DL.opstart);
			slot->d_seg_op = (int32_t*)alloc(sizeof(int32_t) * ((size_t)nno + 1), rc);
			slot->d_seg_op_adj = (int32_t*)alloc(sizeof(int32_t) * ((size_t)nno + 1), rc);
			d.seg_op = slot->d_seg_op;
			if (rc) return rc;
			return finish_upload(ctx, istate, dim, std::move(slot));
		}
		for (void* p : slot->bufs) dev_free(p);
		slot->bufs.clear();
		rc = HDDM_LB_SUCCESS;
	}

	HostLayout L;
	rc = build_layout(dim, nno, dofs, index_i, index_j, ld_index, surplus, ld_surplus, L);
	if (rc) return rc;
	auto up = [&](const void* src, size_t bytes) -> void* {
		if (rc) return nullptr;
		void* p = dev_alloc(bytes, rc);
		if (!p) return nullptr;
		slot->bufs.push_back(p);
		if (bytes && src)
		{
			hipError_t e = hipMemcpy(p, src, bytes, hipMemcpyHostToDevice);
			if (e != hipSuccess) rc = hip_fail(e, "hipMemcpy(upload)", __FILE__, __LINE__);
		}
		return p;
	};
	d.dofs_pad = L.dofs_pad;
	d.nops = (int)L.skip.size() - 1; d.nchunks = L.nchunks; d.max_depth = L.fmax;
	// device copies carry kOpPad sentinels so that read-ahead stays in bounds
	std::vector<Op> ops_pad(L.ops);
	std::vector<uint32_t> skip_pad(L.skip);
	ops_pad.resize((size_t)d.nops + kOpPad, L.ops[d.nops]);
	skip_pad.resize((size_t)d.nops + kOpPad, (uint32_t)d.nops);
	d.ops = (const Op*)up(ops_pad.data(), ops_pad.size() * sizeof(Op));
	d.skip = (const uint32_t*)up(skip_pad.data(), skip_pad.size() * sizeof(uint32_t));
	d.s64 = (const double*)up(L.s64.data(), L.s64.size() * sizeof(double));
	d.s16 = (const uint16_t*)up(L.s16.data(), L.s16.size() * sizeof(uint16_t));
	d.sbf = (const uint16_t*)up(L.sbf.data(), L.sbf.size() * sizeof(uint16_t));
	d.s32 = (const float*)up(L.s32.data(), L.s32.size() * sizeof(float));
	d.s32r = (const float*)up(L.s32r.data(), L.s32r.size() * sizeof(float));
	d.ftab = (const uint32_t*)up(L.ftab.data(), L.ftab.size() * sizeof(uint32_t));
	d.fstride = L.fstride;
	d.iscale = (const double*)up(L.iscale.data(), L.iscale.size() * sizeof(double));
	d.order = (const int32_t*)up(L.order.data(), L.order.size() * sizeof(int32_t));
	d.lomask = (const uint64_t*)up(L.lomask.data(), L.lomask.size() * sizeof(uint64_t));
	d.himask = (const uint64_t*)up(L.himask.data(), L.himask.size() * sizeof(uint64_t));
	slot->d_seg_op = (int32_t*)up(nullptr, sizeof(int32_t) * ((size_t)nno + 1));
	slot->d_seg_op_adj = (int32_t*)up(nullptr, sizeof(int32_t) * ((size_t)nno + 1));
	d.seg_op = slot->d_seg_op;
	slot->shared = std::move(L.shared);
	slot->opstart = std::move(L.opstart);
	if (rc) return rc;
	return finish_upload(ctx, istate, dim, std::move(slot));
}

// Installs a built state in its slot and rebuilds the descriptor tables.
static int finish_upload(hddm_lb_ctx* ctx, int istate, int dim, std::unique_ptr<StateSlot> slot)
{
	slot->loaded = true;
	if ((size_t)istate >= ctx->states.size()) ctx->states.resize(istate + 1);
	ctx->states[istate] = std::move(slot);
	ctx->dim = dim;
	ctx->tables_dirty = true;
	ctx->generation++;
	ctx->tuned.clear();
	// the descriptor tables are built now (not at the next forward), so that
	// the allocation-free forward finds them ready
	return prepare_tables(ctx);
}

extern "C" int hddm_lb_state_array(hddm_lb_ctx* ctx, int istate, int array, void* dst, uint64_t* bytes)
{
	if (!ctx || !bytes) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "state_array: NULL argument");
	if (istate < 0 || (size_t)istate >= ctx->states.size() || !ctx->states[istate] || !ctx->states[istate]->loaded)
		return set_error(HDDM_LB_ERROR_STATE, "state_array: state %d is not loaded", istate);
	const StateDesc& d = ctx->states[istate]->desc;
	const size_t nno = (size_t)d.nno, DP = (size_t)d.dofs_pad, NC = (size_t)d.nchunks, nops = (size_t)d.nops + kOpPad;
	const void* src = nullptr;
	size_t n = 0;
	switch (array)
	{
	case 0: src = d.ops; n = nops * sizeof(Op); break;
	case 1: src = d.skip; n = nops * sizeof(uint32_t); break;
	case 2: src = d.order; n = nno * sizeof(int32_t); break;
	case 3: src = d.s64; n = nno * DP * sizeof(double); break;
	case 4: src = d.s16; n = NC * 2 * DP * kChunk * sizeof(uint16_t); break;
	case 5: src = d.sbf; n = NC * 3 * DP * kChunk * sizeof(uint16_t); break;
	case 6: src = d.s32; n = NC * DP * kChunk * sizeof(float); break;
	case 7: src = d.s32r; n = nno * DP * sizeof(float); break;
	case 8: src = d.iscale; n = DP * sizeof(double); break;
	case 9: src = d.lomask; n = nno * sizeof(uint64_t); break;
	case 10: src = d.himask; n = nno * sizeof(uint64_t); break;
	case 11: src = d.ftab; n = ((nno + 31) & ~(size_t)31) * (size_t)d.fstride * 4 * sizeof(uint32_t); break;
	default: return set_error(HDDM_LB_ERROR_INVALID_VALUE, "state_array: array %d unknown (0 to 11)", array);
	}
	if (dst)
	{
		if (*bytes < n)
			return set_error(HDDM_LB_ERROR_INVALID_VALUE, "state_array: %zu bytes needed, %llu given", n,
				(unsigned long long)*bytes);
		DeviceGuard g(ctx->device);
		HIP_TRY(hipDeviceSynchronize());
		HIP_TRY(hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
	}
	*bytes = n;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_state_info(hddm_lb_ctx* ctx, int istate, int* dim, int* nno, int* dofs, int* max_factors)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "state_info: ctx is NULL");
	if (istate < 0 || (size_t)istate >= ctx->states.size() || !ctx->states[istate] || !ctx->states[istate]->loaded)
		return set_error(HDDM_LB_ERROR_STATE, "state_info: state %d is not loaded", istate);
	const StateDesc& d = ctx->states[istate]->desc;
	if (dim) *dim = d.dim;
	if (nno) *nno = d.nno;
	if (dofs) *dofs = d.dofs;
	if (max_factors) *max_factors = d.max_depth;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_surplus_offset(hddm_lb_ctx* ctx, int istate, int64_t* offset, int64_t* total)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "surplus_offset: ctx is NULL");
	int64_t off = 0, tot = 0;
	bool found = false;
	for (size_t s = 0; s < ctx->states.size(); s++)
	{
		if ((int)s == istate) { off = tot; found = ctx->states[s] && ctx->states[s]->loaded; }
		if (ctx->states[s] && ctx->states[s]->loaded)
			tot += (int64_t)ctx->states[s]->desc.nno * ctx->states[s]->desc.dofs;
	}
	if (istate >= 0 && !found) return set_error(HDDM_LB_ERROR_STATE, "surplus_offset: state %d is not loaded", istate);
	if (offset) *offset = off;
	if (total) *total = tot;
	return HDDM_LB_SUCCESS;
}

namespace hddm_lb {

// Host-pointer calls of at most this many points replay a captured forward.
constexpr int64_t kSmallBatch = 64;

int scratch_wait(hddm_lb_ctx* ctx, hipStream_t st)
{
	if (ctx->scratch_pending && ctx->scratch_stream != st)
		HIP_TRY(hipStreamWaitEvent(st, ctx->scratch_event, 0));
	return HDDM_LB_SUCCESS;
}

int scratch_mark(hddm_lb_ctx* ctx, hipStream_t st)
{
	HIP_TRY(hipEventRecord(ctx->scratch_event, st));
	ctx->scratch_stream = st;
	ctx->scratch_pending = true;
	return HDDM_LB_SUCCESS;
}

// Writes a state's node-segment table for segments of at least `want` nodes
// (0: one segment), boundaries only at top-level runs so each segment's
// first op is at depth 0; returns the segment count.
static int build_segments(const StateSlot& slot, int nno, int64_t want, int32_t* d_table, int& nseg)
{
	std::vector<int32_t> segop;
	segop.push_back(0);
	if (want > 0)
	{
		int64_t last = 0;
		for (int n = 1; n < nno; n++)
			if (n - last >= want && slot.shared[n] == 0) { segop.push_back(slot.opstart[n]); last = n; }
	}
	segop.push_back(slot.opstart[nno]);
	HIP_TRY(hipMemcpy(d_table, segop.data(), sizeof(int32_t) * segop.size(), hipMemcpyHostToDevice));
	nseg = (int)segop.size() - 1;
	return HDDM_LB_SUCCESS;
}

// Uploads the per-state descriptor tables when states or segmenting changed.
int prepare_tables(hddm_lb_ctx* ctx)
{
	const int ns = (int)ctx->states.size();
	if (ns == 0) return set_error(HDDM_LB_ERROR_STATE, "no state is loaded");
	if (!ctx->tables_dirty && ctx->tables_n == ns) return HDDM_LB_SUCCESS;
	// the tables are rewritten with synchronous copies: work still in flight
	// on any stream may read them (this runs only after an upload or an
	// option change)
	HIP_TRY(hipDeviceSynchronize());
	int nloaded = 0;
	for (auto& s : ctx->states) if (s && s->loaded) nloaded++;
	if (!nloaded) return set_error(HDDM_LB_ERROR_STATE, "no state is loaded");

	std::vector<StateDesc> descs(ns), descs_adj(ns);
	std::vector<uint8_t> ok(ns, 0);
	std::vector<std::pair<double, int> > cost;
	ctx->dt_needed = 1;
	ctx->fmax_all = 1;
	ctx->centre_exact_all = true;
	ctx->max_nseg_dp = 0;
	ctx->max_nseg = 1;
	ctx->max_nseg_adj = 1;
	ctx->nseg_adj_total = 0;
	int64_t ds_offset = 0;
	for (int s = 0; s < ns; s++)
	{
		std::memset(&descs[s], 0, sizeof(StateDesc));
		if (!ctx->states[s] || !ctx->states[s]->loaded)
		{
			descs[s].nseg = 1; descs[s].dofs_pad = 32;
			descs_adj[s] = descs[s];
			continue;
		}
		StateSlot& slot = *ctx->states[s];
		StateDesc d = slot.desc;
		// automatic: eighths of states of at least 16k nodes. Segment-major
		// items keep a segment's surplus data L2-resident across the waves
		// and give small batches enough items (measured: D=16 / 60k nodes /
		// 1M points, FP64 sweep 42 -> 32 ms; D=20 / 100k nodes / 64k points,
		// F32 sweep 24 -> 10 ms)
		const int64_t want = ctx->opt_seg_nodes > 0 ? ctx->opt_seg_nodes
			: (d.nno >= 16384 ? ((int64_t)d.nno + 7) / 8 : 0);
		if (slot.seg_len != want)
		{
			int rc = build_segments(slot, d.nno, want, slot.d_seg_op, slot.desc.nseg);
			if (rc) return rc;
			slot.seg_len = want;
		}
		// the wave-specialized adjoint: with owned units, automatic segments of
		// a 32nd of states of at least 16k nodes, so that (state, segment, dS
		// copy) units give every workgroup several units to own; the
		// forward's segments otherwise
		const int64_t want_adj = ctx->opt_seg_nodes > 0 || !ctx->opt_adj_owned ? want
			: (d.nno >= 16384 ? ((int64_t)d.nno + 31) / 32 : 0);
		if (slot.seg_len_adj != want_adj)
		{
			int rc = build_segments(slot, d.nno, want_adj, slot.d_seg_op_adj, slot.nseg_adj);
			if (rc) return rc;
			slot.seg_len_adj = want_adj;
		}
		d = slot.desc;
		d.ds_offset = ds_offset;
		ds_offset += (int64_t)d.nno * d.dofs;
		descs[s] = d;
		descs_adj[s] = d;
		descs_adj[s].nseg = slot.nseg_adj;
		descs_adj[s].seg_op = slot.d_seg_op_adj;
		ctx->max_nseg_adj = std::max(ctx->max_nseg_adj, slot.nseg_adj);
		ctx->nseg_adj_total += slot.nseg_adj;
		ok[s] = 1;
		ctx->dt_needed = std::max(ctx->dt_needed, d.dofs_pad / 32);
		ctx->fmax_all = std::max(ctx->fmax_all, d.max_depth);
		ctx->centre_exact_all = ctx->centre_exact_all && slot.centre_exact;
		if (d.nseg > 1) ctx->max_nseg_dp = std::max<int64_t>(ctx->max_nseg_dp, (int64_t)d.nseg * d.dofs_pad);
		ctx->max_nseg = std::max(ctx->max_nseg, d.nseg);
		cost.push_back(std::make_pair(-(double)d.nops / d.nseg * d.dofs_pad, s));
	}
	std::stable_sort(cost.begin(), cost.end());
	std::vector<int32_t> order;
	for (auto& c : cost) order.push_back(c.second);
	for (int s = 0; s < ns; s++) if (!ok[s]) order.push_back(s);

	if (ctx->tables_cap < ns)
	{
		for (void* p : { (void*)ctx->d_states, (void*)ctx->d_states_adj, (void*)ctx->d_cost_order, (void*)ctx->d_state_ok })
			dev_free(p);
		ctx->d_states = ctx->d_states_adj = nullptr;
		ctx->d_cost_order = nullptr;
		ctx->d_state_ok = nullptr;
		ctx->tables_cap = 0;
		int rc = 0;
		ctx->d_states = (StateDesc*)dev_alloc(sizeof(StateDesc) * ns, rc); if (rc) return rc;
		ctx->d_states_adj = (StateDesc*)dev_alloc(sizeof(StateDesc) * ns, rc); if (rc) return rc;
		ctx->d_cost_order = (int32_t*)dev_alloc(sizeof(int32_t) * ns, rc); if (rc) return rc;
		ctx->d_state_ok = (uint8_t*)dev_alloc(ns, rc); if (rc) return rc;
		ctx->tables_cap = ns;
	}
	HIP_TRY(hipMemcpy(ctx->d_states, descs.data(), sizeof(StateDesc) * ns, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(ctx->d_states_adj, descs_adj.data(), sizeof(StateDesc) * ns, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(ctx->d_cost_order, order.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice));
	HIP_TRY(hipMemcpy(ctx->d_state_ok, ok.data(), ns, hipMemcpyHostToDevice));
	ctx->tables_n = ns;
	ctx->tables_dirty = false;
	return HDDM_LB_SUCCESS;
}

template <typename T>
static int grow(T*& p, size_t& cap, size_t need)
{
	if (need <= cap && p) return HDDM_LB_SUCCESS;
	// a buffer being replaced may still be read by work in flight on any stream
	if (p) HIP_TRY(hipDeviceSynchronize());
	dev_free(p);
	p = nullptr;
	cap = 0;
	int rc = 0;
	size_t n = std::max<size_t>(need, 1);
	p = (T*)dev_alloc(sizeof(T) * n, rc);
	if (rc) return rc;
	cap = n;
	return HDDM_LB_SUCCESS;
}

// Pinned host buffer of at least `need` elements (hipHostMalloc; never
// hipHostRegister).
template <typename T>
static int grow_pinned(T*& p, size_t& cap, size_t need)
{
	if (need <= cap && p) return HDDM_LB_SUCCESS;
	if (p) { HIP_TRY(hipDeviceSynchronize()); (void)hipHostFree(p); }
	p = nullptr;
	cap = 0;
	const size_t n = std::max<size_t>(need, 64);
	void* q = nullptr;
	hipError_t e = hipHostMalloc(&q, sizeof(T) * n, hipHostMallocDefault);
	if (e != hipSuccess)
	{
		(void)hipGetLastError();
		return set_error(HDDM_LB_ERROR_OUT_OF_MEMORY, "hipHostMalloc(%zu): %s", sizeof(T) * n, hipGetErrorString(e));
	}
	p = (T*)q;
	cap = n;
	return HDDM_LB_SUCCESS;
}

int bucket_bits(int64_t npoints, int nstates, int sort)
{
	if (!sort) return 0;
	int64_t per = npoints / std::max(1, nstates) / 16;
	int b = 0;
	while (b < 16 && ((int64_t)1 << (b + 1)) <= per) b++;
	while (b > 0 && ((int64_t)(nstates + 1) << b) > (1 << 18)) b--;
	return b;
}

// Allocates scratch for npoints; must run outside graph capture.
int prepare_scratch(hddm_lb_ctx* ctx, Scratch& sc, int64_t npoints)
{
	int rc = prepare_tables(ctx);
	if (rc) return rc;
	const int ns = ctx->tables_n;
	const int bb = bucket_bits(npoints, ns, ctx->opt_sort);
	const size_t nb = (size_t)(ns + 1) << bb;
	if ((rc = grow(sc.bkey, sc.cap_bkey, (size_t)npoints))) return rc;
	if ((rc = grow(sc.perm, sc.cap_perm, (size_t)npoints))) return rc;
	if ((rc = grow(sc.hist, sc.cap_hist, nb))) return rc;
	if ((rc = grow(sc.cursor, sc.cap_cursor, nb))) return rc;
	if ((rc = grow(sc.plans, sc.cap_plans, (size_t)ns))) return rc;
	if ((rc = grow(sc.hdr, sc.cap_hdr, 1))) return rc;
	if (ctx->max_nseg_dp > 0)
		if ((rc = grow(sc.partial, sc.cap_partial, (size_t)npoints * (size_t)ctx->max_nseg_dp))) return rc;
	return HDDM_LB_SUCCESS;
}

TuneConfig default_config(hddm_lb_ctx* ctx, int prec)
{
	TuneConfig c;
	c.waves = ctx->opt_waves ? ctx->opt_waves : 4;
	c.walk = ctx->opt_walk == 2 ? kWalkLds : kWalkSmem;
	int slab = ctx->opt_slab;
	// automatic: a four-bank (D > 32) MFMA sweep runs one wave per SIMD
	// either way, so two dof tiles per pass halve the walks of wide states
	// (measured: D=59, 128 dofs, F32 670 -> 432 ms); narrower ones keep
	// their occupancy with one tile
	if (!slab) slab = xbanks_for(ctx->dim) == 4 && ctx->dt_needed >= 2 ? 2 : 1;
	if (prec == HDDM_LB_FP64) slab = 1;
	c.slab = slab;
	// automatic: FP64 sweeps of wide states (D > 32) of two or more dof tiles
	// are walked once by a walker wave feeding one consumer wave per tile
	// (measured: D=59, 128 dofs, 2^18 points: FP64 962 -> 731 ms; the split
	// and F32 paths are faster with every wave walking its own block, 408 vs
	// 505 ms for f16x2)
	c.ws = ctx->opt_form == 2
		|| (ctx->opt_form == 0 && prec == HDDM_LB_FP64 && xbanks_for(ctx->dim) == 4 && ctx->dt_needed >= 2);
	c.nc = ctx->dt_needed >= 2 ? 4 : 2;
	return c;
}

// Enqueues the whole forward sequence on st. No allocation, no host sync.
int enqueue_forward(hddm_lb_ctx* ctx, Scratch& sc, int64_t npoints, const double* x, int64_t ldx,
	const int32_t* sid, double* y, int64_t ldy, int prec, hipStream_t st, const TuneConfig& cfg, bool timed)
{
	hipStreamCaptureStatus cap_status = hipStreamCaptureStatusNone;
	HIP_TRY(hipStreamIsCapturing(st, &cap_status));
	const bool in_capture = cap_status != hipStreamCaptureStatusNone;
	const int ns = ctx->tables_n;
	const int bb = bucket_bits(npoints, ns, ctx->opt_sort);
	const int64_t nb = (int64_t)(ns + 1) << bb;
	HIP_TRY(launch_fill_u32(sc.hist, nb, 0u, st));
	HIP_TRY(launch_bucket_key(npoints, x, ldx, sid, ctx->d_state_ok, ns, ctx->dim, bb, sc.bkey, sc.hist, st));
	HIP_TRY(launch_scan_plan(sc.hist, sc.cursor, nb, bb, ns, ctx->d_states, ctx->d_cost_order,
		sc.plans, sc.hdr, !cfg.ws && cfg.walk == kWalkLds ? cfg.waves : 1, 0, st));
	HIP_TRY(launch_scatter(npoints, sc.bkey, sc.cursor, sc.perm, st));

	const void* kern = cfg.ws ? sweep_ws_kernel_ptr(prec, xbanks_for(ctx->dim), ctx->fmax_all, cfg.nc)
		: sweep_kernel_ptr(prec, cfg.slab, xbanks_for(ctx->dim), ctx->fmax_all, cfg.walk);
	if (!kern) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: no sweep kernel for precision %d, dim %d", prec, ctx->dim);
	const int waves = cfg.ws ? cfg.nc + 1 : cfg.waves;
	const size_t lds = cfg.ws ? sweep_ws_lds_bytes(prec, ctx->fmax_all) : sweep_lds_bytes(prec, cfg.waves, ctx->fmax_all, cfg.walk);
	HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	int per_cu = 0;
	HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, waves * kWave, lds));
	if (per_cu < 1)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "forward: sweep kernel does not fit a CU (%d waves)", waves);
	const int grid = per_cu * ctx->num_cus;
	// dof tiles per launch: the WS sweep's consumers cover 4 tiles (nc 4) or 1
	const int tiles = cfg.ws ? (cfg.nc == 4 ? 4 : 1) : cfg.slab;

	SweepArgs a;
	std::memset(&a, 0, sizeof(a));
	a.states = ctx->d_states; a.plans = sc.plans; a.cost_order = ctx->d_cost_order; a.hdr = sc.hdr;
	a.perm = sc.perm; a.x = x; a.ldx = ldx; a.y = y; a.ldy = ldy; a.partial = sc.partial;
	a.nstates = ns; a.dim = ctx->dim;
	a.status = ctx->d_status;
	if (ctx->opt_stats && !in_capture)
	{
		if (!ctx->d_stats)
		{
			int rc = 0;
			ctx->d_stats = (unsigned long long*)dev_alloc(hddm_lb_ctx::kStatCount * sizeof(unsigned long long), rc);
			if (rc) return rc;
		}
		HIP_TRY(launch_fill_u32((uint32_t*)ctx->d_stats, 2 * hddm_lb_ctx::kStatCount, 0u, st));
		a.stats = ctx->d_stats;
	}
	const int nslab = (ctx->dt_needed + tiles - 1) / tiles;
	auto& ring = ctx->tring[0];
	const int slot = (int)(ring.calls % hddm_lb_ctx::kTimingRing);
	if (timed) HIP_TRY(hipEventRecord(ring.e0[slot], st));
	for (int sl = 0; sl < nslab; sl++)
	{
		a.dt_base = sl * tiles;
		if (sl > 0) HIP_TRY(launch_reset_counter(sc.hdr, st));
		HIP_TRY(launch_sweep(kern, grid, waves, lds, a, st));
	}
	if (timed)
	{
		HIP_TRY(hipEventRecord(ring.e1[slot], st));
		ring.calls++;
	}
	if (ctx->max_nseg_dp > 0)
		HIP_TRY(launch_reduce(ctx->d_states, sc.plans, sc.perm, sc.partial, y, ldy, ns, st));
	return HDDM_LB_SUCCESS;
}

int check_forward_args(hddm_lb_ctx* ctx, int64_t npoints, const double* x, int64_t ldx,
	double* y, int64_t ldy, int prec)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: ctx is NULL");
	if (npoints < 0 || npoints > ((int64_t)1 << 31) - 1)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: npoints %lld out of range", (long long)npoints);
	if (prec < HDDM_LB_FP64 || prec > HDDM_LB_F32)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: unknown precision %d", prec);
	if (ctx->dim < 1) return set_error(HDDM_LB_ERROR_STATE, "forward: no state is loaded");
	if (npoints && (!x || !y)) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: x and y must not be NULL");
	if (ldx < ctx->dim) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: ldx %lld < dim %d", (long long)ldx, ctx->dim);
	int maxdofs = 0;
	for (auto& s : ctx->states) if (s && s->loaded) maxdofs = std::max(maxdofs, s->desc.dofs);
	if (ldy < maxdofs) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "forward: ldy %lld < dofs %d", (long long)ldy, maxdofs);
	return HDDM_LB_SUCCESS;
}

} // namespace hddm_lb

extern "C" int hddm_lb_forward(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream)
{
	int rc = check_forward_args(ctx, npoints, x, ldx, y, ldy, precision);
	if (rc) return rc;
	if (npoints == 0) return HDDM_LB_SUCCESS;
	DeviceGuard g(ctx->device);
	hipStream_t st = (hipStream_t)stream;
	if ((rc = prepare_scratch(ctx, ctx->scratch, npoints))) return rc;
	if ((rc = scratch_wait(ctx, st))) return rc;
	TuneConfig cfg = default_config(ctx, precision);
	if (ctx->opt_autotune && (rc = autotune(ctx, npoints, x, ldx, state_ids, y, ldy, precision, st, cfg)))
		return rc;
	ctx->last_timed = ctx->opt_timing;
	ctx->last_cfg = cfg;
	rc = enqueue_forward(ctx, ctx->scratch, npoints, x, ldx, state_ids, y, ldy, precision, st, cfg, ctx->opt_timing);
	if (rc) return rc;
	return scratch_mark(ctx, st);
}

extern "C" int hddm_lb_last_stats_n(hddm_lb_ctx* ctx, uint64_t* stats, int n)
{
	if (!ctx || !stats) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_stats: NULL argument");
	if (n < 1 || n > hddm_lb_ctx::kStatCount)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_stats: %d counters requested, 1 to %d available", n, hddm_lb_ctx::kStatCount);
	if (!ctx->d_stats) return set_error(HDDM_LB_ERROR_STATE, "last_stats: HDDM_LB_OPT_STATS was not set for a forward");
	DeviceGuard g(ctx->device);
	HIP_TRY(hipDeviceSynchronize());
	HIP_TRY(hipMemcpy(stats, ctx->d_stats, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_last_config(hddm_lb_ctx* ctx, int* config)
{
	if (!ctx || !config) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_config: NULL argument");
	const TuneConfig& c = ctx->last_cfg;
	config[0] = c.ws ? 2 : 1;
	config[1] = c.ws ? c.nc + 1 : c.waves;
	config[2] = c.ws ? (c.nc == 4 ? 4 : 1) : c.slab;
	config[3] = (int)ctx->tune_runs;
	config[4] = (int)ctx->tune_hits;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_last_adjoint_config(hddm_lb_ctx* ctx, int* config)
{
	if (!ctx || !config) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_adjoint_config: NULL argument");
	for (int i = 0; i < 4; i++) config[i] = ctx->last_adj[i];
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_last_adjoint_stats(hddm_lb_ctx* ctx, uint64_t* stats, int n)
{
	if (!ctx || !stats) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_adjoint_stats: NULL argument");
	if (n < 1 || n > hddm_lb_ctx::kAdjStatCount)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "last_adjoint_stats: %d counters requested, 1 to %d available", n,
			hddm_lb_ctx::kAdjStatCount);
	if (!ctx->d_adj_stats) return set_error(HDDM_LB_ERROR_STATE, "last_adjoint_stats: HDDM_LB_OPT_STATS was not set for an adjoint");
	DeviceGuard g(ctx->device);
	HIP_TRY(hipDeviceSynchronize());
	HIP_TRY(hipMemcpy(stats, ctx->d_adj_stats, (size_t)n * sizeof(uint64_t), hipMemcpyDeviceToHost));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_last_stats(hddm_lb_ctx* ctx, uint64_t* stats)
{
	return hddm_lb_last_stats_n(ctx, stats, 4);
}

extern "C" double hddm_lb_last_sweep_ms(hddm_lb_ctx* ctx)
{
	if (!ctx || !ctx->last_timed || ctx->tring[0].calls == 0) return -1.0;
	auto& ring = ctx->tring[0];
	DeviceGuard g(ctx->device);
	const int slot = (int)((ring.calls - 1) % hddm_lb_ctx::kTimingRing);
	if (hipEventSynchronize(ring.e1[slot]) != hipSuccess) { (void)hipGetLastError(); return -1.0; }
	float ms = -1.0f;
	if (hipEventElapsedTime(&ms, ring.e0[slot], ring.e1[slot]) != hipSuccess) { (void)hipGetLastError(); return -1.0; }
	return (double)ms;
}

extern "C" int hddm_lb_kernel_time_summary(hddm_lb_ctx* ctx, int kind, int64_t* ncalls, double* total_ms,
	double* min_ms, double* max_ms)
{
	if (!ctx || !ncalls || !total_ms) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "kernel_time_summary: NULL argument");
	if (kind < 0 || kind > 1) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "kernel_time_summary: kind %d is not 0 or 1", kind);
	DeviceGuard g(ctx->device);
	auto& ring = ctx->tring[kind];
	const int64_t first = std::max(ring.read, ring.calls - hddm_lb_ctx::kTimingRing);
	double tot = 0.0, mn = 0.0, mx = 0.0;
	for (int64_t c = first; c < ring.calls; c++)
	{
		const int slot = (int)(c % hddm_lb_ctx::kTimingRing);
		HIP_TRY(hipEventSynchronize(ring.e1[slot]));
		float ms = 0.0f;
		HIP_TRY(hipEventElapsedTime(&ms, ring.e0[slot], ring.e1[slot]));
		tot += ms;
		mn = c == first ? ms : std::min(mn, (double)ms);
		mx = c == first ? ms : std::max(mx, (double)ms);
	}
	*ncalls = ring.calls - first;
	*total_ms = tot;
	if (min_ms) *min_ms = mn;
	if (max_ms) *max_ms = mx;
	ring.read = ring.calls;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_sweep_time_summary(hddm_lb_ctx* ctx, int64_t* ncalls, double* total_ms,
	double* min_ms, double* max_ms)
{
	return hddm_lb_kernel_time_summary(ctx, 0, ncalls, total_ms, min_ms, max_ms);
}

struct hddm_lb_plan
{
	hddm_lb_ctx* ctx;
	uint64_t generation;
	hddm_lb::Scratch scratch;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	~hddm_lb_plan()
	{
		if (exec) (void)hipGraphExecDestroy(exec);
		if (graph) (void)hipGraphDestroy(graph);
	}
};

// The staged forward of host buffers; the caller holds ctx->run_mu.
static int forward_host_run(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision)
{
	int rc = HDDM_LB_SUCCESS;
	DeviceGuard g(ctx->device);
	if (!ctx->own_stream) HIP_TRY(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
	const bool small = npoints <= kSmallBatch;
	// small batches size the device staging for kSmallBatch points, so the
	// captured forwards that point at it stay valid until a row stride grows
	const size_t nx = (size_t)(small ? kSmallBatch : npoints) * ldx;
	const size_t ny = (size_t)(small ? kSmallBatch : npoints) * ldy;
	const size_t ns = (size_t)(small ? kSmallBatch : npoints);
	if (nx > ctx->cap_xstage || ny > ctx->cap_ystage || (state_ids && ns > ctx->cap_sidstage))
	{
		for (auto& kv : ctx->small_plans) hddm_lb_plan_destroy(kv.second);
		ctx->small_plans.clear();
	}
	if ((rc = grow(ctx->d_xstage, ctx->cap_xstage, nx))) return rc;
	if ((rc = grow(ctx->d_ystage, ctx->cap_ystage, ny))) return rc;
	if (state_ids && (rc = grow(ctx->d_sidstage, ctx->cap_sidstage, ns))) return rc;
	if ((rc = grow_pinned(ctx->h_xstage, ctx->cap_hx, (size_t)npoints * ldx))) return rc;
	if ((rc = grow_pinned(ctx->h_ystage, ctx->cap_hy, (size_t)npoints * ldy))) return rc;
	if (state_ids && (rc = grow_pinned(ctx->h_sidstage, ctx->cap_hsid, (size_t)npoints))) return rc;
	hipStream_t st = ctx->own_stream;
	// pageable caller memory -> pinned staging on the host, one DMA each way
	std::memcpy(ctx->h_xstage, x, sizeof(double) * npoints * ldx);
	if (state_ids) std::memcpy(ctx->h_sidstage, state_ids, sizeof(int32_t) * npoints);
	// rows of points with an invalid state stay as the caller passed them
	std::memcpy(ctx->h_ystage, y, sizeof(double) * npoints * ldy);
	if ((rc = scratch_wait(ctx, st))) return rc;
	HIP_TRY(hipMemcpyAsync(ctx->d_xstage, ctx->h_xstage, sizeof(double) * npoints * ldx, hipMemcpyHostToDevice, st));
	if (state_ids)
		HIP_TRY(hipMemcpyAsync(ctx->d_sidstage, ctx->h_sidstage, sizeof(int32_t) * npoints, hipMemcpyHostToDevice, st));
	HIP_TRY(hipMemcpyAsync(ctx->d_ystage, ctx->h_ystage, sizeof(double) * npoints * ldy, hipMemcpyHostToDevice, st));
	if (small)
	{
		auto key = std::make_tuple(npoints, ldx, ldy, state_ids ? 1 : 0, precision);
		auto it = ctx->small_plans.find(key);
		if (it != ctx->small_plans.end() && it->second->generation != ctx->generation)
		{
			hddm_lb_plan_destroy(it->second);
			ctx->small_plans.erase(it);
			it = ctx->small_plans.end();
		}
		if (it == ctx->small_plans.end())
		{
			hddm_lb_plan* plan = nullptr;
			if ((rc = hddm_lb_plan_create(ctx, npoints, ctx->d_xstage, ldx, state_ids ? ctx->d_sidstage : nullptr,
					ctx->d_ystage, ldy, precision, st, &plan)))
				return rc;
			it = ctx->small_plans.emplace(key, plan).first;
		}
		if ((rc = hddm_lb_plan_launch(it->second, st))) return rc;
	}
	else
	{
		rc = hddm_lb_forward(ctx, npoints, ctx->d_xstage, ldx, state_ids ? ctx->d_sidstage : nullptr,
			ctx->d_ystage, ldy, precision, st);
		if (rc) return rc;
	}
	HIP_TRY(hipMemcpyAsync(ctx->h_ystage, ctx->d_ystage, sizeof(double) * npoints * ldy, hipMemcpyDeviceToHost, st));
	HIP_TRY(hipStreamSynchronize(st));
	std::memcpy(y, ctx->h_ystage, sizeof(double) * npoints * ldy);
	return HDDM_LB_SUCCESS;
}

// Requests taken into one coalesced batch, in points.
constexpr int64_t kCoalesceMax = 4096;

// Runs queued requests of one precision as one forward: rows packed with
// ldx = dim and ldy = the widest state's dofs (the callers' y rows are packed
// too, so columns and rows the forward leaves alone come back unchanged).
static int run_coalesced(hddm_lb_ctx* ctx, const std::vector<HostRequest*>& batch)
{
	std::lock_guard<std::mutex> run(ctx->run_mu);
	if (batch.size() == 1)
	{
		const HostRequest* r = batch[0];
		return forward_host_run(ctx, r->n, r->x, r->ldx, r->sid, r->y, r->ldy, r->prec);
	}
	int maxdofs = 0;
	for (auto& s : ctx->states) if (s && s->loaded) maxdofs = std::max(maxdofs, s->desc.dofs);
	const int dim = ctx->dim;
	int64_t tot = 0;
	for (const HostRequest* r : batch) tot += r->n;
	std::vector<double> xs((size_t)tot * dim), ys((size_t)tot * maxdofs);
	std::vector<int32_t> sid((size_t)tot);
	int64_t row = 0;
	for (const HostRequest* r : batch)
	{
		const int64_t w = std::min<int64_t>(r->ldy, maxdofs);
		for (int64_t p = 0; p < r->n; p++, row++)
		{
			std::memcpy(&xs[(size_t)row * dim], r->x + p * r->ldx, sizeof(double) * dim);
			std::memcpy(&ys[(size_t)row * maxdofs], r->y + p * r->ldy, sizeof(double) * w);
			sid[(size_t)row] = r->sid ? r->sid[p] : 0;
		}
	}
	const int rc = forward_host_run(ctx, tot, xs.data(), dim, sid.data(), ys.data(), maxdofs, batch[0]->prec);
	if (rc) return rc;
	row = 0;
	for (const HostRequest* r : batch)
	{
		const int64_t w = std::min<int64_t>(r->ldy, maxdofs);
		for (int64_t p = 0; p < r->n; p++, row++)
			std::memcpy(r->y + p * r->ldy, &ys[(size_t)row * maxdofs], sizeof(double) * w);
	}
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_forward_host(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision)
{
	int rc = check_forward_args(ctx, npoints, x, ldx, y, ldy, precision);
	if (rc) return rc;
	if (npoints == 0) return HDDM_LB_SUCCESS;
	if (npoints > kSmallBatch || !ctx->opt_coalesce)
	{
		std::lock_guard<std::mutex> run(ctx->run_mu);
		return forward_host_run(ctx, npoints, x, ldx, state_ids, y, ldy, precision);
	}
	// small call: queue it; whichever waiting thread finds no batch running
	// takes every queued request of its precision (up to kCoalesceMax points)
	// and runs them as one forward
	HostRequest me{ npoints, x, ldx, state_ids, y, ldy, precision, HDDM_LB_SUCCESS, false };
	std::unique_lock<std::mutex> lk(ctx->comb_mu);
	ctx->comb_queue.push_back(&me);
	while (!me.done)
	{
		if (ctx->comb_busy) { ctx->comb_cv.wait(lk); continue; }
		ctx->comb_busy = true;
		std::vector<HostRequest*> batch;
		int64_t tot = 0;
		for (auto it = ctx->comb_queue.begin(); it != ctx->comb_queue.end();)
		{
			if ((*it)->prec == precision && tot + (*it)->n <= kCoalesceMax)
			{
				tot += (*it)->n;
				batch.push_back(*it);
				it = ctx->comb_queue.erase(it);
			}
			else ++it;
		}
		lk.unlock();
		const int brc = batch.empty() ? HDDM_LB_SUCCESS : run_coalesced(ctx, batch);
		lk.lock();
		for (HostRequest* r : batch) { r->rc = brc; r->done = true; }
		ctx->comb_requests += (int64_t)batch.size();
		ctx->comb_batches += batch.empty() ? 0 : 1;
		ctx->comb_busy = false;
		ctx->comb_cv.notify_all();
	}
	return me.rc;
}

extern "C" int hddm_lb_host_call_stats(hddm_lb_ctx* ctx, int64_t* requests, int64_t* batches)
{
	if (!ctx || !requests || !batches) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "host_call_stats: NULL argument");
	std::lock_guard<std::mutex> lk(ctx->comb_mu);
	*requests = ctx->comb_requests;
	*batches = ctx->comb_batches;
	return HDDM_LB_SUCCESS;
}

// ---- hipGraph plans ----


extern "C" int hddm_lb_plan_create(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream, hddm_lb_plan** out)
{
	if (!out) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plan_create: plan is NULL");
	*out = nullptr;
	int rc = check_forward_args(ctx, npoints, x, ldx, y, ldy, precision);
	if (rc) return rc;
	if (npoints < 1) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plan_create: npoints must be >= 1");
	DeviceGuard g(ctx->device);
	if ((rc = prepare_scratch(ctx, ctx->scratch, npoints))) return rc;
	TuneConfig cfg = default_config(ctx, precision);
	if (ctx->opt_autotune)
	{
		if ((rc = scratch_wait(ctx, (hipStream_t)stream))) return rc;
		if ((rc = autotune(ctx, npoints, x, ldx, state_ids, y, ldy, precision, (hipStream_t)stream, cfg))) return rc;
		if ((rc = scratch_mark(ctx, (hipStream_t)stream))) return rc;
	}
	std::unique_ptr<hddm_lb_plan> plan(new hddm_lb_plan);
	plan->ctx = ctx;
	plan->generation = ctx->generation;
	if ((rc = prepare_scratch(ctx, plan->scratch, npoints))) return rc;
	HIP_TRY(hipDeviceSynchronize());
	hipStream_t cap;
	HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
	hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
	if (e != hipSuccess) { (void)hipStreamDestroy(cap); return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__); }
	rc = enqueue_forward(ctx, plan->scratch, npoints, x, ldx, state_ids, y, ldy, precision, cap, cfg, false);
	hipGraph_t graph = nullptr;
	e = hipStreamEndCapture(cap, &graph);
	(void)hipStreamDestroy(cap);
	if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
	if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
	plan->graph = graph;
	HIP_TRY(hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0));
	*out = plan.release();
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_plan_launch(hddm_lb_plan* plan, void* stream)
{
	if (!plan) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plan_launch: plan is NULL");
	if (plan->generation != plan->ctx->generation)
		return set_error(HDDM_LB_ERROR_STATE, "plan_launch: the context's states changed after the plan was created");
	DeviceGuard g(plan->ctx->device);
	HIP_TRY(hipGraphLaunch(plan->exec, (hipStream_t)stream));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_plan_destroy(hddm_lb_plan* plan)
{
	if (plan)
	{
		DeviceGuard g(plan->ctx->device);
		delete plan;
	}
	return HDDM_LB_SUCCESS;
}
