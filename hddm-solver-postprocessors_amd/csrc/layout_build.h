// Host-side construction of a state's device layout (CPU only, testable without a GPU).
#pragma once

#include "layout.h"

#include <cstdint>
#include <vector>

namespace hddm_lb {

struct HostLayout
{
	int dim = 0, nno = 0, dofs = 0, dofs_pad = 0, fmax = 0, rec_words = 0, nchunks = 0;
	std::vector<uint32_t> rec;        // [nno * rec_words]
	std::vector<Factor> fac;          // unique factors, ordered by (dim, i, j)
	std::vector<double> s64;          // [nno * dofs_pad]
	std::vector<uint16_t> s16;        // [nchunks][2][dofs_pad][16]
	std::vector<uint16_t> sbf;        // [nchunks][3][dofs_pad][16]
	std::vector<float> s32;           // [nchunks][dofs_pad][2][8]
	std::vector<double> iscale;       // [dofs_pad]
	std::vector<int32_t> order;       // trie position -> uploaded node index
	std::vector<uint64_t> lomask, himask;  // [nno]
};

// Builds the prefix-trie layout. Returns an HDDM_LB_* status.
int build_layout(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, HostLayout& out);

} // namespace hddm_lb
