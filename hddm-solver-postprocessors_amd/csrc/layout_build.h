// Host-side construction of a state's device layout (CPU only, testable without a GPU).
#pragma once

#include "layout.h"

#include <cstdint>
#include <functional>
#include <vector>

namespace hddm_lb {

struct HostLayout
{
	int dim = 0, nno = 0, dofs = 0, dofs_pad = 0, fmax = 0, nchunks = 0;
	std::vector<Op> ops;              // [nops + 1], trie order, sentinel last
	std::vector<uint32_t> skip;       // [nops]
	std::vector<int32_t> opstart;     // [nno + 1] first op of each node
	std::vector<uint8_t> shared;      // [nno] factors shared with the previous node
	std::vector<double> s64;          // [nno * dofs_pad]
	std::vector<uint16_t> s16;        // [nchunks][2][dofs_pad][16]
	std::vector<uint16_t> sbf;        // [nchunks][3][dofs_pad][16]
	std::vector<float> s32;           // [nchunks][dofs_pad][2][8]
	std::vector<float> s32r;          // [nno * dofs_pad] FP32 rows (adjoint's g . S products)
	int fstride = 0;                  // factor slots per node in ftab: 4, 8 or 16
	std::vector<uint32_t> ftab;       // [nno][fstride][4] factors of each node, trie order
	std::vector<double> iscale;       // [dofs_pad]
	std::vector<int32_t> order;       // trie position -> uploaded node index
	std::vector<uint64_t> lomask, himask;  // [nno]
};

// Segment boundaries for a target length: nodes where a new top-level run
// starts (shared prefix 0), so every segment's first op is at depth 0.
std::vector<int32_t> segment_nodes(const HostLayout& L, int64_t seg_len);

// Builds the prefix-trie layout. Returns an HDDM_LB_* status.
int build_layout(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, HostLayout& out);

// The same layout built on the current device (layout_device.hip): device
// arrays equal, bit for bit, to build_layout's, plus the host copies of the
// shared prefix lengths and first ops that segment planning reads.
struct DeviceLayout
{
	int nops = 0, fmax = 0, fstride = 0, dofs_pad = 0, nchunks = 0;
	Op* ops = nullptr;            // [nops + kOpPad]
	uint32_t* skip = nullptr;     // [nops + kOpPad]
	double* s64 = nullptr;
	uint16_t* s16 = nullptr;
	uint16_t* sbf = nullptr;
	float* s32 = nullptr;
	float* s32r = nullptr;
	double* iscale = nullptr;
	int32_t* order = nullptr;
	uint64_t* lomask = nullptr;
	uint64_t* himask = nullptr;
	uint32_t* ftab = nullptr;
	std::vector<uint8_t> shared;  // host
	std::vector<int32_t> opstart; // host, nno + 1
};
// Device memory for the arrays of a DeviceLayout, owned by the caller; on
// failure it sets rc and returns NULL.
typedef std::function<void*(size_t, int&)> DeviceAlloc;
// Returned by build_layout_device for inputs it leaves to build_layout
// (out-of-range indices, more than 15 factors, too many nodes or ops), which
// either builds them or reports the error.
constexpr int kLayoutHostCheck = -1;
int build_layout_device(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, const DeviceAlloc& alloc, DeviceLayout& out);

} // namespace hddm_lb
