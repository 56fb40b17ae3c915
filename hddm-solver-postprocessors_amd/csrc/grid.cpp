// Seeded sparse-grid generator for synthetic inputs (no reference counterpart:
// the reference reads grids written by the HDDM solver).
//
// 1-D hierarchy of the reference's hat basis (LinearBasis/gold/src/Data.cpp:474-494):
//   level 1: one node, i = 0, j = 0 (phi == 1)
//   level 2: boundary nodes x = 0, 1: i = 2, j in {0, 2}
//   level l >= 3: i = 2^(l-1), odd j, node at x = j / i
// Children: level 1 -> both level-2 nodes; (2, 0) -> (3, 1); (2, 2) -> (3, 3);
// (l, j) -> (l+1, 2j-1), (l+1, 2j+1) for l >= 3.
#include "hddm_lb.h"
#include "errors.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

struct hddm_lb_grid
{
	int dim;
	// Active factors of each node, packed (dim << 16) | (level << 8) | j, ascending dim.
	std::vector<std::vector<uint32_t> > nodes;
};

namespace {

struct SplitMix64
{
	uint64_t s;
	explicit SplitMix64(uint64_t seed) : s(seed) { }
	uint64_t next()
	{
		uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
		z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
		z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
		return z ^ (z >> 31);
	}
	double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint32_t pack(int d, int l, int j) { return ((uint32_t)d << 16) | ((uint32_t)l << 8) | (uint32_t)j; }
inline int fdim(uint32_t f) { return (int)(f >> 16); }
inline int flev(uint32_t f) { return (int)((f >> 8) & 0xff); }
inline int fj(uint32_t f) { return (int)(f & 0xff); }

int children1d(int l, int j, int* cl, int* cj)
{
	if (l == 1) { cl[0] = 2; cj[0] = 0; cl[1] = 2; cj[1] = 2; return 2; }
	if (l == 2) { cl[0] = 3; cj[0] = (j == 0) ? 1 : 3; return 1; }
	cl[0] = l + 1; cj[0] = 2 * j - 1; cl[1] = l + 1; cj[1] = 2 * j + 1; return 2;
}

std::string key_of(const std::vector<uint32_t>& n)
{
	return std::string(reinterpret_cast<const char*>(n.data()), n.size() * sizeof(uint32_t));
}

int level_sum(const std::vector<uint32_t>& n)
{
	int s = 0;
	for (uint32_t f : n) s += flev(f) - 1;
	return s;
}

struct Builder
{
	int dim, maxsum;
	std::vector<std::vector<uint32_t> >& nodes;
	std::vector<int> sums;
	std::vector<char> refined;
	std::unordered_map<std::string, int> index;

	Builder(int dim_, int maxsum_, std::vector<std::vector<uint32_t> >& nodes_) :
		dim(dim_), maxsum(maxsum_), nodes(nodes_) { }

	bool add(std::vector<uint32_t>&& n)
	{
		std::string k = key_of(n);
		if (index.count(k)) return false;
		index.emplace(std::move(k), (int)nodes.size());
		sums.push_back(level_sum(n));
		refined.push_back(0);
		nodes.push_back(std::move(n));
		return true;
	}

	// Adds every child of node idx (in every dim) whose level sum is <= limit.
	void refine(int idx, int limit)
	{
		refined[idx] = 1;
		const std::vector<uint32_t> base = nodes[idx];
		for (int d = 0; d < dim; d++)
		{
			int l = 1, j = 0, pos = -1;
			for (size_t t = 0; t < base.size(); t++)
				if (fdim(base[t]) == d) { l = flev(base[t]); j = fj(base[t]); pos = (int)t; break; }
			if (sums[idx] + 1 > limit) continue;
			int cl[2], cj[2];
			int nc = children1d(l, j, cl, cj);
			for (int c = 0; c < nc; c++)
			{
				std::vector<uint32_t> child = base;
				uint32_t f = pack(d, cl[c], cj[c]);
				if (pos >= 0) child[pos] = f;
				else
				{
					child.push_back(f);
					std::sort(child.begin(), child.end());
				}
				add(std::move(child));
			}
		}
	}
};

} // namespace

extern "C" int hddm_lb_grid_generate(int dim, int start_sum, int max_level, double p_refine,
	uint64_t seed, int64_t target_nno, hddm_lb_grid** grid)
{
	using namespace hddm_lb;
	if (!grid) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "grid_generate: grid is NULL");
	*grid = nullptr;
	if (dim < 1 || dim > 65535 || start_sum < 0 || max_level < 1 || max_level > 8 ||
		start_sum > max_level - 1 || !(p_refine >= 0.0 && p_refine <= 1.0))
		return set_error(HDDM_LB_ERROR_INVALID_VALUE,
			"grid_generate: need 1 <= dim <= 65535, 1 <= max_level <= 8, 0 <= start_sum <= max_level-1, 0 <= p <= 1");

	hddm_lb_grid* g = new hddm_lb_grid;
	g->dim = dim;
	Builder b(dim, max_level - 1, g->nodes);
	b.add(std::vector<uint32_t>());
	// Regular grid: breadth-first closure under children with level sum <= start_sum.
	for (size_t idx = 0; idx < g->nodes.size(); idx++)
		b.refine((int)idx, start_sum);
	// Nodes below the start sum already have all their children.
	for (size_t idx = 0; idx < g->nodes.size(); idx++)
		b.refined[idx] = (b.sums[idx] < start_sum) ? 1 : 0;

	SplitMix64 rng(seed);
	const int limit = max_level - 1;
	// With p = 0 no node is ever refined: the grid is the regular start grid
	// (and the passes below would never run out of candidates).
	bool done = p_refine <= 0.0 || (target_nno > 0 && (int64_t)g->nodes.size() >= target_nno);
	while (!done)
	{
		bool any = false;
		size_t n0 = g->nodes.size();
		for (size_t idx = 0; idx < n0 && !done; idx++)
		{
			if (b.refined[idx] || b.sums[idx] >= limit) continue;
			any = true;
			if (rng.uniform() < p_refine)
				b.refine((int)idx, limit);
			if (target_nno > 0 && (int64_t)g->nodes.size() >= target_nno) done = true;
		}
		if (!any) break;
	}
	*grid = g;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_grid_nno(const hddm_lb_grid* grid, int64_t* nno)
{
	using namespace hddm_lb;
	if (!grid || !nno) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "grid_nno: NULL argument");
	*nno = (int64_t)grid->nodes.size();
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_grid_index(const hddm_lb_grid* grid, int32_t* index_i, int32_t* index_j)
{
	using namespace hddm_lb;
	if (!grid || !index_i || !index_j) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "grid_index: NULL argument");
	const int dim = grid->dim;
	for (size_t k = 0; k < grid->nodes.size(); k++)
	{
		int32_t* ik = index_i + k * dim;
		int32_t* jk = index_j + k * dim;
		std::memset(ik, 0, sizeof(int32_t) * dim);
		std::memset(jk, 0, sizeof(int32_t) * dim);
		for (uint32_t f : grid->nodes[k])
		{
			ik[fdim(f)] = 1 << (flev(f) - 1);
			jk[fdim(f)] = fj(f);
		}
	}
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_grid_levels(const hddm_lb_grid* grid, int32_t* levels, int32_t* indices)
{
	using namespace hddm_lb;
	if (!grid || !levels || !indices) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "grid_levels: NULL argument");
	const int dim = grid->dim;
	for (size_t k = 0; k < grid->nodes.size(); k++)
	{
		int32_t* lk = levels + k * dim;
		int32_t* xk = indices + k * dim;
		for (int d = 0; d < dim; d++) { lk[d] = 1; xk[d] = 1; }
		// The uncompressed .plt stores index = j + 1 (LinearBasis/gold/src/Data.cpp:487-492).
		for (uint32_t f : grid->nodes[k]) { lk[fdim(f)] = flev(f); xk[fdim(f)] = fj(f) + 1; }
	}
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_grid_free(hddm_lb_grid* grid)
{
	delete grid;
	return HDDM_LB_SUCCESS;
}
