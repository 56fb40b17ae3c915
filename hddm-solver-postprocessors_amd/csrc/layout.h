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

// One factor evaluation of the prefix-trie walk ("op"), 16 bytes, read by
// one scalar load. The ops of all nodes are stored in trie order; node n
// contributes one op per factor it does not share with node n-1.
//   dim    twice the dimension of x the factor reads (bits 0-7; the index of
//          its first 32-bit register in a bank of FP64 coordinates)
//          | (skip[op] - op) << 8
//   ctrl   depth t (bits 0-3) | last factor of its node (bit 4)
//          | relation to the previous op in the stream (bits 5-6)
//          | the op has children (bit 7) | node index << 8
//   ihi    high word of the FP64 value of i (the low word is 0)
//   njhi   high word of the FP64 value of -j (the low word is 0)
// The walk computes P[t+1] = P[t] * max(0, 1 - |x_dim * i - j|). When P[t+1]
// is zero on every lane, the run of nodes sharing the node's first t+1
// factors is skipped: skip[op] is the first op after that run.
// Relation (precomputed so the walk needs no depth compares): 0 the op is
// the first child of the previous op (its P[t] is the previous product),
// 1 a sibling (same P[t] as the previous op), 2 otherwise (P[t] is read from
// the path stack). After a skip of more than one op the target's stream
// predecessor is deeper, so its relation is 2 and the stack holds P[t].
// Children: P[t+1] is needed later only when the next op descends from this
// one, so only such ops store it on the stack.
// The device copies of ops and skip carry kOpPad sentinels after the stream.
struct Op
{
	uint32_t dim, ctrl, ihi, njhi;
};
constexpr uint32_t kOpLast = 16u;
constexpr int kOpRelShift = 5;
constexpr uint32_t kOpKids = 128u;
constexpr int kOpNodeShift = 8;
constexpr uint32_t kOpDimMask = 0xffu;
constexpr int kOpSkipShift = 8;
constexpr int kOpPad = 4;

// Dimension of x read by an op with dim word w (constexpr: host and device).
constexpr uint32_t op_dim(uint32_t w) { return (w & kOpDimMask) >> 1; }

struct StateDesc
{
	int dim, nno, dofs, dofs_pad;   // dofs_pad: dofs rounded up to 32
	int nops, nchunks;
	int nseg, max_depth;            // node segments (work-item split), max factors per node
	const Op* ops;                  // [nops + kOpPad]
	const uint32_t* skip;           // [nops + kOpPad]
	const int32_t* seg_op;          // [nseg + 1] first op of each segment
	const double* s64;              // [nno * dofs_pad] FP64 rows, trie order
	const uint16_t* s16;            // [nchunks][2][dofs_pad][16] f16 hi, lo (scaled)
	const uint16_t* sbf;            // [nchunks][3][dofs_pad][16] bf16 terms
	const float* s32;               // [nchunks][dofs_pad][2][8] f32
	const float* s32r;              // [nno * dofs_pad] f32 rows, trie order
	// Factors of every node (trie order), fstride slots of 16 bytes each:
	// {twice the dim, high word of i, high word of -j / i (the centre
	// negated), 1}, then {0, 0, 0, 0} padding (phi == 1 at u = 0, no
	// derivative). u = (x - j / i) i equals the op's fma(x, i, -j) bit for
	// bit: i is a power of two.
	const uint32_t* ftab;           // [nno rounded up to 32][fstride][4]
	int64_t fstride;
	const double* iscale;           // [dofs_pad] 2^-sigma_c undoing the f16 scaling
	const int32_t* order;           // trie position -> uploaded node index
	const uint64_t* lomask;         // [nno] dims whose support lies in [0, 1/2]
	const uint64_t* himask;         // [nno] dims whose support lies in [1/2, 1]
	int64_t ds_offset;              // first element of the state's rows in an adjoint dsurplus
};

// Per-state plan of one forward call, written on the device by the plan kernel.
struct StatePlan
{
	int64_t start;         // first sorted position of the state's points
	int64_t count;         // number of points of the state
	int64_t nblk;          // 64-point blocks
	int64_t item_base;     // first work item (items ordered by state cost)
	int64_t partial_base;  // element offset in the partial buffer (split states)
	int64_t dx_base;       // element offset in the adjoint's dx partial buffer (split states)
};

struct PlanHeader
{
	int64_t nitems;
	unsigned long long counter;   // work-queue head, reset by the plan kernel
	int64_t partial_elems;
	int64_t pad;
};

} // namespace hddm_lb
