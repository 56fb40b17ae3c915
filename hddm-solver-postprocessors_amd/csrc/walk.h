// The per-wave prefix-trie walk shared by the forward sweep and the adjoint's
// dS kernel (gfx950 only): the hand-scheduled scalar-load walk for one, two or
// four coordinate banks and the walk state (path stack, chunk buffer, op
// position).
#pragma once

#include "sweep_common.h"
#include "kernels.h"

namespace hddm_lb {

typedef double d16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t lds_offset(const void* p)
{
	return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// Hand-scheduled scalar-load walk (D <= 64). Walks ops from byte offset qb
// until the first surviving node of a chunk other than `chunk` (returned in
// node and val, not stored; qb is then past its op) or until qb reaches qe
// (node = -1).
// Bounds, shown on the host side: every relative skip is >= 1, so qb grows
// each step and the loop ends; a successor, the op two ahead and a skip
// target lie at most two ops past the last op of the stream, inside its
// kOpPad sentinels; the depth t <= 14 and P has a row for depth t + 1 (the
// stage holds 8 rows when every node has at most 7 factors, else 16), also
// for the read issued ahead for the op after a segment's last (an op of any
// state, or a sentinel of depth 0); the
// register index 2 * dim stays inside the pinned coordinate banks.
// Per op: the parent product comes from the LDS path stack (P[t],
// lane-strided rows of 512 bytes at pl) and the op's product is stored as
// P[t+1], so P[k] is the product of the first k factors of the current trie
// path. The read of P[t] is issued at the end of the previous step (or after
// a skip target is loaded, or on entry), as soon as the op's depth is known;
// a wave's LDS operations complete in order, so it sees the P[t+1] its parent
// step has just stored; the hat's fma reads the coordinate straight from the bank as its
// first source under GPR indexing (s_set_gpr_idx_on, index = the op's dim
// field, twice the dimension). A wave-wide zero
// product skips the op's run (the skip target and its successor are loaded
// only when the target is not the successor); a live last factor emits its
// node into the chunk buffer at bl.
// Ops rotate through three SGPR sets s[76:79], s[80:83], s[84:87]: while op
// q is evaluated, op q+1 is already in flight (issued one step earlier) and
// op q+2 is issued right after the step's one wait, so a scalar-cache miss
// has a whole step to return. The three steps follow each other in the
// code: the common path takes one branch per three ops, prunes and the
// emission of a node of a new chunk branch out. The op handed back is in
// s[76:79].
// Fixed registers: the FP64 i as s[94:95] (low word 0), x from v64 (one,
// two or four banks of 32 VGPRs), temporaries s88-s90 and eight VGPRs after
// x (HW_*; the low word of -j is kept 0 in HW_B0). GPR indexing writes m0,
// which the compiler reserves and may hold live across the statement: it is
// copied to s91 on entry and restored at the one exit (L_done), so the
// statement leaves m0 as it found it.
// The chunk buffer holds doubles (rows of 512 bytes) or, for the F32 path,
// floats (rows of 256 bytes): HW_BROW (the row offset of a chunk position,
// a shift for rows of 256 or 512 bytes, else a multiply: padded rows) and
// HW_BST (the store).
// Issues the read of P[t] for the op whose ctrl word is C1 (into HW_C, its
// address in HW_D).
#define HDDM_WALK_PREP(C1)                                                             \
	"s_lshl_b32 s88, " C1 ", 9\n"                                                      \
	"s_and_b32 s88, s88, 0x1e00\n"                                                     \
	"v_add_u32 " HW_D ", s88, %[pl]\n"                                                     \
	"ds_read_b64 " HW_C ", " HW_D "\n"

#define HDDM_WALK_MAIN(X, X0, X1, X2, X3, Y1, ZQ, NEXT)                                \
	"L_top" X "_%=:\n"                                                                 \
	"s_mov_b32 s95, " X2 "\n"                                                          \
	"v_mov_b32 " HW_B1 ", " X3 "\n"                                                          \
	"s_set_gpr_idx_on " X0 ", gpr_idx(SRC0)\n"                                         \
	"v_fma_f64 " HW_A ", v[64:65], s[94:95], " HW_B "\n"                               \
	"s_set_gpr_idx_off\n"                                                              \
	"v_add_f64 " HW_A ", -|" HW_A "|, 1.0\n"                                           \
	"v_max_f64 " HW_A ", " HW_A ", 0\n"                                                \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	"s_load_dwordx4 " ZQ ", %[ob], %[qb] offset:0x20\n"                                \
	"v_mul_f64 " HW_A ", " HW_C ", " HW_A "\n"                                       \
	"ds_write_b64 " HW_D ", " HW_A " offset:512\n"                                         \
	"s_add_u32 %[nops], %[nops], 1\n"                                                  \
	"v_cmp_neq_f64_e32 vcc, 0, " HW_A "\n"                                             \
	"s_cbranch_vccz L_prune" X "_%=\n"                                                 \
	"s_bitcmp0_b32 " X1 ", 4\n"                                                        \
	"s_cbranch_scc1 L_adv" X "_%=\n"                                                   \
	"s_lshr_b32 s89, " X1 ", %[cs]\n"                                                  \
	"s_cmp_lg_u32 s89, %[chunk]\n"                                                     \
	"s_cbranch_scc1 L_new" X "_%=\n"                                                   \
	"s_bfe_u32 s89, " X1 ", %[cf]\n"                                                   \
	"s_bitset1_b32 %[mask], s89\n"                                                     \
	HW_BROW                                                                            \
	"v_add_u32 " HW_E ", s89, %[bl]\n"                                                 \
	HW_BST                                                                             \
	"L_adv" X "_%=:\n"                                                                 \
	HDDM_WALK_PREP(Y1)                                                                 \
	"s_add_u32 %[qb], %[qb], 16\n"                                                     \
	"s_cmp_lt_u32 %[qb], %[qe]\n"                                                      \
	NEXT

// continue with step Y: falls through when Y's step follows, else branches back
#define HDDM_WALK_FALL(Y) "s_cbranch_scc0 L_end" Y "_%=\n"
#define HDDM_WALK_LOOP(Y) "s_cbranch_scc1 L_top" Y "_%=\n" "s_branch L_end" Y "_%=\n"

#define HDDM_WALK_SIDE(X, X0, X1, Y, Y1, YQ, ZQ)                                       \
	"L_prune" X "_%=:\n"                                                               \
	"s_add_u32 %[nprunes], %[nprunes], 1\n"                                            \
	"s_lshr_b32 s90, " X0 ", 8\n"                                                      \
	"s_cmp_eq_u32 s90, 1\n"                                                            \
	"s_cbranch_scc1 L_adv" X "_%=\n"                                                   \
	"s_lshl_b32 s90, s90, 4\n"                                                         \
	"s_add_u32 %[qb], %[qb], s90\n"                                                    \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	"s_load_dwordx4 " YQ ", %[ob], %[qb]\n"                                            \
	"s_load_dwordx4 " ZQ ", %[ob], %[qb] offset:0x10\n"                                \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	HDDM_WALK_PREP(Y1)                                                                 \
	"s_cmp_lt_u32 %[qb], %[qe]\n"                                                      \
	"s_cbranch_scc1 L_top" Y "_%=\n"                                                   \
	"s_branch L_end" Y "_%=\n"                                                         \
	"L_new" X "_%=:\n"                                                                 \
	"s_add_u32 %[qb], %[qb], 16\n"                                                     \
	"s_lshr_b32 %[node], " X1 ", 8\n"                                                  \
	"v_mov_b64_e32 %[val], " HW_A "\n"                                                 \
	"s_branch L_exit" Y "_%=\n"

#define HDDM_WALK_ASM                                                                  \
	"s_mov_b32 s91, m0\n"                                                              \
	"v_mov_b32 " HW_B0 ", 0\n"                                                               \
	"s_mov_b32 s94, 0\n"                                                               \
	"s_load_dwordx4 s[80:83], %[ob], %[qb] offset:0x10\n"                              \
	HDDM_WALK_PREP("s77")                                                              \
	HDDM_WALK_MAIN("0", "s76", "s77", "s78", "s79", "s81", "s[84:87]", HDDM_WALK_FALL("1")) \
	HDDM_WALK_MAIN("1", "s80", "s81", "s82", "s83", "s85", "s[76:79]", HDDM_WALK_FALL("2")) \
	HDDM_WALK_MAIN("2", "s84", "s85", "s86", "s87", "s77", "s[80:83]", HDDM_WALK_LOOP("0")) \
	HDDM_WALK_SIDE("0", "s76", "s77", "1", "s81", "s[80:83]", "s[84:87]")              \
	HDDM_WALK_SIDE("1", "s80", "s81", "2", "s85", "s[84:87]", "s[76:79]")              \
	HDDM_WALK_SIDE("2", "s84", "s85", "0", "s77", "s[76:79]", "s[80:83]")              \
	"L_end0_%=:\n"                                                                     \
	"s_mov_b32 %[node], -1\n"                                                          \
	"L_exit0_%=:\n"                                                                    \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	"s_branch L_done_%=\n"                                                             \
	"L_end1_%=:\n"                                                                     \
	"s_mov_b32 %[node], -1\n"                                                          \
	"L_exit1_%=:\n"                                                                    \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	"s_mov_b64 s[76:77], s[80:81]\n"                                                   \
	"s_mov_b64 s[78:79], s[82:83]\n"                                                   \
	"s_branch L_done_%=\n"                                                             \
	"L_end2_%=:\n"                                                                     \
	"s_mov_b32 %[node], -1\n"                                                          \
	"L_exit2_%=:\n"                                                                    \
	"s_waitcnt lgkmcnt(0)\n"                                                           \
	"s_mov_b64 s[76:77], s[84:85]\n"                                                   \
	"s_mov_b64 s[78:79], s[86:87]\n"                                                   \
	"L_done_%=:\n"                                                                     \
	"s_mov_b32 m0, s91\n"

#define HDDM_WALK_OUTPUTS                                                              \
	: [qb] "+s"(qb), [mask] "+s"(mask), [nops] "+s"(nops), [nprunes] "+s"(nprunes),    \
	  [node] "=&s"(node), [val] "=&v"(val), "+{s[76:79]}"(opv)
#define HDDM_WALK_INPUTS                                                               \
	: [ob] "s"(ob), [qe] "s"(qe), [chunk] "s"(chunk), [pl] "v"(pl), [bl] "v"(bl),      \
	  [cs] "i"(8 + CB), [cf] "i"((CB << 16) | 8), [rb] "i"(RB)
#define HDDM_WALK_SCLOBBERS                                                            \
	"s80", "s81", "s82", "s83", "s84", "s85", "s86", "s87", "s88", "s89", "s90",       \
	"s91", "s94", "s95", "vcc", "scc", "memory"

// One coordinate bank (D <= 16): x in v[64:95], temporaries v[96:103].
// BF: the chunk buffer holds floats, else doubles; RB: bytes per row (256 or
// 512 for 64 values, more for padded rows).
template <int CB, bool BF, int RB>
__device__ __forceinline__ int walk_op_stream(const char* ob, uint32_t& qb, uint32_t qe, int chunk,
	u4_t& opv, uint32_t& mask, uint32_t& nops, uint32_t& nprunes, uint32_t pl, uint32_t bl,
	const d16_t& xv, double& val)
{
	int node;
#define HW_A "v[96:97]"
#define HW_A0 "v96"
#define HW_A1 "v97"
#define HW_B "v[98:99]"
#define HW_B0 "v98"
#define HW_B1 "v99"
#define HW_C "v[100:101]"
#define HW_D "v102"
#define HW_E "v103"
	if constexpr (BF && RB == 256)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 8\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv)
			: HDDM_WALK_SCLOBBERS, "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (!BF && RB == 512)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 9\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv)
			: HDDM_WALK_SCLOBBERS, "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (BF)
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv)
			: HDDM_WALK_SCLOBBERS, "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103");
#undef HW_BROW
#undef HW_BST
	}
	else
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv)
			: HDDM_WALK_SCLOBBERS, "v96", "v97", "v98", "v99", "v100", "v101", "v102", "v103");
#undef HW_BROW
#undef HW_BST
	}
#undef HW_A
#undef HW_A0
#undef HW_A1
#undef HW_B
#undef HW_B0
#undef HW_B1
#undef HW_C
#undef HW_D
#undef HW_E
	return node;
}

// Two coordinate banks (16 < D <= 32): x in v[64:127] (the indexed source
// reaches dims 16-31 through the same index), temporaries v[128:135].
template <int CB, bool BF, int RB>
__device__ __forceinline__ int walk_op_stream2(const char* ob, uint32_t& qb, uint32_t qe, int chunk,
	u4_t& opv, uint32_t& mask, uint32_t& nops, uint32_t& nprunes, uint32_t pl, uint32_t bl,
	const d16_t& xv0, const d16_t& xv1, double& val)
{
	int node;
#define HW_A "v[128:129]"
#define HW_A0 "v128"
#define HW_A1 "v129"
#define HW_B "v[130:131]"
#define HW_B0 "v130"
#define HW_B1 "v131"
#define HW_C "v[132:133]"
#define HW_D "v134"
#define HW_E "v135"
	if constexpr (BF && RB == 256)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 8\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1)
			: HDDM_WALK_SCLOBBERS, "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (!BF && RB == 512)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 9\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1)
			: HDDM_WALK_SCLOBBERS, "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (BF)
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1)
			: HDDM_WALK_SCLOBBERS, "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135");
#undef HW_BROW
#undef HW_BST
	}
	else
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1)
			: HDDM_WALK_SCLOBBERS, "v128", "v129", "v130", "v131", "v132", "v133", "v134", "v135");
#undef HW_BROW
#undef HW_BST
	}
#undef HW_A
#undef HW_A0
#undef HW_A1
#undef HW_B
#undef HW_B0
#undef HW_B1
#undef HW_C
#undef HW_D
#undef HW_E
	return node;
}

// Four coordinate banks (32 < D <= 64): x in v[64:191], temporaries v[192:199].
template <int CB, bool BF, int RB>
__device__ __forceinline__ int walk_op_stream4(const char* ob, uint32_t& qb, uint32_t qe, int chunk,
	u4_t& opv, uint32_t& mask, uint32_t& nops, uint32_t& nprunes, uint32_t pl, uint32_t bl,
	const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3, double& val)
{
	int node;
#define HW_A "v[192:193]"
#define HW_A0 "v192"
#define HW_A1 "v193"
#define HW_B "v[194:195]"
#define HW_B0 "v194"
#define HW_B1 "v195"
#define HW_C "v[196:197]"
#define HW_D "v198"
#define HW_E "v199"
	if constexpr (BF && RB == 256)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 8\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1), "{v[128:159]}"(xv2), "{v[160:191]}"(xv3)
			: HDDM_WALK_SCLOBBERS, "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (!BF && RB == 512)
	{
#define HW_BROW "s_lshl_b32 s89, s89, 9\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1), "{v[128:159]}"(xv2), "{v[160:191]}"(xv3)
			: HDDM_WALK_SCLOBBERS, "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199");
#undef HW_BROW
#undef HW_BST
	}
	else if constexpr (BF)
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "v_cvt_f32_f64 " HW_A0 ", " HW_A "\n" "ds_write_b32 " HW_E ", " HW_A0 "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1), "{v[128:159]}"(xv2), "{v[160:191]}"(xv3)
			: HDDM_WALK_SCLOBBERS, "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199");
#undef HW_BROW
#undef HW_BST
	}
	else
	{
#define HW_BROW "s_mul_i32 s89, s89, %[rb]\n"
#define HW_BST "ds_write_b64 " HW_E ", " HW_A "\n"
		asm volatile(HDDM_WALK_ASM HDDM_WALK_OUTPUTS HDDM_WALK_INPUTS, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1), "{v[128:159]}"(xv2), "{v[160:191]}"(xv3)
			: HDDM_WALK_SCLOBBERS, "v192", "v193", "v194", "v195", "v196", "v197", "v198", "v199");
#undef HW_BROW
#undef HW_BST
	}
#undef HW_A
#undef HW_A0
#undef HW_A1
#undef HW_B
#undef HW_B0
#undef HW_B1
#undef HW_C
#undef HW_D
#undef HW_E
	return node;
}

#undef HDDM_WALK_OUTPUTS
#undef HDDM_WALK_INPUTS
#undef HDDM_WALK_SCLOBBERS
#undef HDDM_WALK_ASM
#undef HDDM_WALK_MAIN
#undef HDDM_WALK_SIDE
#undef HDDM_WALK_FALL
#undef HDDM_WALK_LOOP

// Walk state of one wave. P[t] (the product of the first t factors of the
// current trie path) lives in the wave's LDS stage; the two most recent
// products stay in registers, so only a move up the trie reads LDS.
// WALK selects where ops come from:
//   kWalkSmem  scalar loads from the state's op stream, the sequential
//              successor read one op ahead; waves are independent
//   kWalkLds   the workgroup's LDS tile (wave-uniform vector reads)
// ROW: elements per row of the chunk buffer (64, or more for rows padded
// against LDS bank conflicts of column reads).
template <int XB, typename BT, int WALK, int PM, bool XSEL, int CB, int ROW = kWave>
struct Walker
{
	static constexpr int kRB = ROW * (int)sizeof(BT);
	static constexpr int kC = 1 << CB;        // nodes per chunk
	static constexpr int kCM = kC - 1;
	double* P;          // [PM][64] in LDS
	BT* B;              // [chunk nodes][64] in LDS: basis values of the current chunk
	double base_prev;   // P[last_t]
	double cur;         // P[last_t + 1]
	int last_t;
	int q, q1;
	int chunk;          // chunk whose positions B holds (-1: none)
	uint32_t mask;      // positions of B written in this chunk
	uint32_t nops, nprunes;  // counts for measurement
	const Op* gops;     // kWalkSmem: the op stream and skips
	const uint32_t* gskip;
	Op op;              // kWalkSmem: op q, in SGPRs

	__device__ __forceinline__ void init(int q0, int qend, double* Pl, BT* Bl, int lane,
		const Op* ops, const uint32_t* skip)
	{
		q = q0;
		q1 = qend;
		P = Pl;
		B = Bl;
		P[lane] = 1.0;
		base_prev = 1.0;
		cur = 1.0;
		last_t = -1;
		chunk = -1;
		mask = 0;
		nops = 0;
		nprunes = 0;
		gops = ops;
		gskip = skip;
		if constexpr (WALK == kWalkSmem) op = ldc_op(ops + q);
	}

	__device__ __forceinline__ double pget(int t, int lane) { return P[t * kWave + lane]; }

	__device__ __forceinline__ void pset(int t, double v, int lane) { P[t * kWave + lane] = v; }

	// One factor at depth t: the new prefix product.
	__device__ __forceinline__ double step(int t, uint32_t dim, uint32_t ihi, uint32_t njhi, int lane,
		const double (&xb0)[16], const double (&xb1)[16], const double (&xb2)[16], const double (&xb3)[16])
	{
		double bt;
		if (t == last_t + 1) bt = cur;
		else if (t == last_t) bt = base_prev;
		else bt = pget(t, lane);
		const double tt = fma(xget<XB, XSEL>(xb0, xb1, xb2, xb3, dim), f64_from_high(ihi), f64_from_high(njhi));
		const double p = bt * fmax(1.0 - fabs(tt), 0.0);
		pset(t + 1, p, lane);
		base_prev = bt;
		cur = p;
		last_t = t;
		nops++;
		return p;
	}

	// Walks ops below tend (the end of the resident tile or of the segment),
	// writing basis values of nodes of `chunk` into B. Stops at the first
	// surviving node of another chunk, returned in (node, val) and not yet
	// stored (returns true), or when q reaches tend (returns false).
	__device__ __forceinline__ bool run(const Op* tops, const uint32_t* tskip, int tb, int tend, int lane,
		const double (&xb0)[16], const double (&xb1)[16], const double (&xb2)[16], const double (&xb3)[16],
		int& node, double& val)
	{
		if constexpr (WALK == kWalkSmem)
		{
			uint32_t qb = (uint32_t)q * (uint32_t)sizeof(Op);
			const uint32_t qe = (uint32_t)tend * (uint32_t)sizeof(Op);
			if (qb >= qe) return false;
			u4_t opv = { op.dim, op.ctrl, op.ihi, op.njhi };
			d16_t xv0, xv1, xv2, xv3;
#pragma unroll
			for (int i = 0; i < 16; i++) { xv0[i] = xb0[i]; xv1[i] = xb1[i]; xv2[i] = xb2[i]; xv3[i] = xb3[i]; }
			double v = 0.0;
			const uint32_t pl = lds_offset(P) + (uint32_t)lane * 8u;
			const uint32_t bl = lds_offset(B) + (uint32_t)lane * (uint32_t)sizeof(BT);
			int n;
			if constexpr (XB == 1)
				n = walk_op_stream<CB, sizeof(BT) == 4, kRB>((const char*)gops, qb, qe, chunk, opv, mask, nops, nprunes,
					pl, bl, xv0, v);
			else if constexpr (XB == 2)
				n = walk_op_stream2<CB, sizeof(BT) == 4, kRB>((const char*)gops, qb, qe, chunk, opv, mask, nops, nprunes,
					pl, bl, xv0, xv1, v);
			else
				n = walk_op_stream4<CB, sizeof(BT) == 4, kRB>((const char*)gops, qb, qe, chunk, opv, mask, nops, nprunes,
					pl, bl, xv0, xv1, xv2, xv3, v);
			op.dim = opv[0]; op.ctrl = opv[1]; op.ihi = opv[2]; op.njhi = opv[3];
			q = (int)(qb / (uint32_t)sizeof(Op));
			if (n < 0) return false;
			node = n;
			val = v;
			return true;
		}
		else
		{
			while (q < tend)
			{
				const int sl = q - tb;
				const u4_t ov = *(const u4_t*)(tops + sl);
				const uint32_t dim = op_dim(__builtin_amdgcn_readfirstlane(ov[0]));
				const uint32_t ctrl = __builtin_amdgcn_readfirstlane(ov[1]);
				const uint32_t ihi = __builtin_amdgcn_readfirstlane(ov[2]);
				const uint32_t njhi = __builtin_amdgcn_readfirstlane(ov[3]);
				const double p = step((int)(ctrl & 15u), dim, ihi, njhi, lane, xb0, xb1, xb2, xb3);
				if (!wave_any(p != 0.0))
				{
					nprunes++;
					q = (int)__builtin_amdgcn_readfirstlane(tskip[sl]);
					continue;
				}
				q++;
				if (ctrl & kOpLast)
				{
					const int n = (int)(ctrl >> kOpNodeShift);
					if ((n >> CB) != chunk) { node = n; val = p; return true; }
					B[(n & kCM) * ROW + lane] = (BT)p;
					mask |= 1u << (n & kCM);
				}
			}
			return false;
		}
	}

	__device__ __forceinline__ void start_chunk(int node, double val, int lane)
	{
		chunk = node >> CB;
		mask = 1u << (node & kCM);
		B[(node & kCM) * ROW + lane] = (BT)val;
	}
};

} // namespace hddm_lb
