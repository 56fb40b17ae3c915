// Device-side data layout of one uploaded state and of one forward call.
// Shared by the host library (capi.cpp, state.cpp) and the HIP kernels.
#pragma once

#include <stdint.h>

namespace hddm_lb {

// Points handled by one wave: one point per lane.
constexpr int kWave = 64;
// Nodes per MFMA chunk (the K extent of v_mfma_f32_32x32x16_*).
constexpr int kChunk = 16;
// Bytes of x staged per dim for one wave: 64 lanes x FP64.
constexpr int kXRowBytes = kWave * 8;
// Bits of the spatial sort key below the state id.
constexpr int kMortonBits = 40;

// One unique hat factor phi(x_dim; i, j) = 1 - |x_dim * i - j|.
struct alignas(32) Factor
{
	double i;        // 0 or 2^(l-1), any integer in general
	double negj;     // -j
	uint32_t xoff;   // dim * kXRowBytes: byte offset of x_dim in a wave's LDS stage
	uint32_t dim;
	uint32_t lomask_bit;  // reserved
	uint32_t pad;
};

// Node record in prefix-trie order, stride rec_words 32-bit words:
//   word 0: F (active factors) | s (factors shared with the previous node) << 8
//   word 1: unused
//   words 2 + 2t, 3 + 2t (t < fmax): factor id of depth t, and the index of the
//           first node after the subtree that shares this node's first t+1 factors.
struct StateDesc
{
	int dim, nno, dofs, dofs_pad;   // dofs_pad: dofs rounded up to 32
	int fmax, rec_words, nchunks;
	int seg_len, nseg;              // node segments: [seg*seg_len, (seg+1)*seg_len)
	int pad0;
	const uint32_t* rec;            // [nno * rec_words]
	const Factor* fac;              // unique factors
	const double* s64;              // [nno * dofs_pad] FP64 rows, trie order
	const uint16_t* s16;            // [nchunks][2][dofs_pad][16] f16 hi, lo (scaled)
	const uint16_t* sbf;            // [nchunks][3][dofs_pad][16] bf16 terms
	const float* s32;               // [nchunks][dofs_pad][2][8] f32
	const double* iscale;           // [dofs_pad] 2^-sigma_c undoing the f16 scaling
	const int32_t* order;           // trie position -> uploaded node index
	const uint64_t* lomask;         // [nno] dims whose support lies in [0, 1/2]
	const uint64_t* himask;         // [nno] dims whose support lies in [1/2, 1]
};

// Per-state plan of one forward call, written on the device by the plan kernel.
struct StatePlan
{
	int64_t start;         // first sorted position of the state's points
	int64_t count;         // number of points of the state
	int64_t nblk;          // 64-point blocks
	int64_t item_base;     // first work item (items ordered by state cost)
	int64_t partial_base;  // element offset in the partial buffer (split states)
};

struct PlanHeader
{
	int64_t nitems;
	unsigned long long counter;   // work-queue head, reset by the plan kernel
	int64_t partial_elems;
	int64_t pad;
};

} // namespace hddm_lb
