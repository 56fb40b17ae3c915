// Host-side construction of a state's device layout (CPU only, testable without a GPU).
#pragma once

#include "layout.h"

#include <cstdint>
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

} // namespace hddm_lb
