// .plt surplus files: the on-disk format the HDDM solver writes and the
// reference's DataDense::load reads (LinearBasis/gold/src/Data.cpp:13-595).
//
// Compressed layout (LinearBasis/gold/src/Data.cpp:349-363, 50-190, 525-572):
//   "compressed" | i32 dim | i32 nno | i32 dofs | i32 level
//   "index"   | u32 nnz | nnz x {u16 i, u16 j} | (nno+1) x T row-length deltas | nnz x T dim
//   "surplus" | u32 nnz | nnz x f64           | (nno+1) x U row-length deltas | nnz x U dof
//   with T (U) = u8 if dim (dofs) <= 255, u16 if <= 65535, else u32.
//   (i, j) are stored already transformed (i = 0 or 2^(l-1)).
// Uncompressed layout (LinearBasis/gold/src/Data.cpp:364-398, 449-524):
//   i32 dim | i32 nno | i32 dofs | i32 level, then per node 2*dim ints
//   (levels, indices) and dofs doubles; binary, or the same numbers as text when
//   the first int read in binary differs from nagents. Level l maps to
//   i = 2 << (l - 2) and j = index - 1, with j = 0 wherever i == 0.
//
// Differences from the reference, all on inputs the reference treats as fatal
// or undefined: errors are returned instead of aborting; level 1 maps to i = 0
// explicitly (the reference computes 2 << -1); levels < 1 and short reads are
// format errors.
#include "hddm_lb.h"
#include "errors.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace {

using hddm_lb::set_error;

struct File
{
	FILE* f;
	explicit File(FILE* f_) : f(f_) { }
	~File() { if (f) fclose(f); }
};

bool read_exact(FILE* f, void* dst, size_t bytes)
{
	return bytes == 0 || fread(dst, 1, bytes, f) == bytes;
}

bool read_marker(FILE* f, const char* marker)
{
	size_t n = strlen(marker);
	char buf[16] = { 0 };
	return read_exact(f, buf, n) && memcmp(buf, marker, n) == 0;
}

template <typename T>
bool read_deltas(FILE* f, int nno, std::vector<int64_t>& ia)
{
	std::vector<T> t(nno + 1);
	if (!read_exact(f, t.data(), sizeof(T) * t.size())) return false;
	ia.resize(nno + 1);
	ia[0] = t[0];
	for (int r = 1; r <= nno; r++) ia[r] = ia[r - 1] + (int64_t)t[r];
	return true;
}

template <typename T>
bool read_cols(FILE* f, size_t nnz, std::vector<uint32_t>& ja)
{
	std::vector<T> t(nnz);
	if (!read_exact(f, t.data(), sizeof(T) * nnz)) return false;
	ja.assign(t.begin(), t.end());
	return true;
}

int width_for(int n)
{
	if (n <= (int)std::numeric_limits<uint8_t>::max()) return 1;
	if (n <= (int)std::numeric_limits<uint16_t>::max()) return 2;
	return 4;
}

bool read_csr_tail(FILE* f, int nno, int ncols, size_t nnz, std::vector<int64_t>& ia, std::vector<uint32_t>& ja)
{
	int w = width_for(ncols);
	bool ok = (w == 1) ? read_deltas<uint8_t>(f, nno, ia) :
		(w == 2) ? read_deltas<uint16_t>(f, nno, ia) : read_deltas<uint32_t>(f, nno, ia);
	if (!ok) return false;
	return (w == 1) ? read_cols<uint8_t>(f, nnz, ja) :
		(w == 2) ? read_cols<uint16_t>(f, nnz, ja) : read_cols<uint32_t>(f, nnz, ja);
}

int check_csr(const char* path, const char* what, int nno, int ncols, size_t nnz,
	const std::vector<int64_t>& ia, const std::vector<uint32_t>& ja)
{
	if (ia[0] != 0)
		return set_error(HDDM_LB_ERROR_FORMAT, "%s: %s IA[0] must be 0", path, what);
	for (int r = 1; r <= nno; r++)
		if (ia[r] < ia[r - 1])
			return set_error(HDDM_LB_ERROR_FORMAT, "%s: %s IA must be non-decreasing (row %d)", path, what, r);
	if ((size_t)ia[nno] != nnz)
		return set_error(HDDM_LB_ERROR_FORMAT, "%s: %s row lengths sum to %lld, expected %zu",
			path, what, (long long)ia[nno], nnz);
	for (size_t e = 0; e < nnz; e++)
		if ((int64_t)ja[e] >= ncols)
			return set_error(HDDM_LB_ERROR_FORMAT, "%s: %s column %u out of range [0, %d)", path, what, ja[e], ncols);
	return HDDM_LB_SUCCESS;
}

template <typename T>
void write_deltas(FILE* f, const std::vector<int64_t>& rowlen)
{
	std::vector<T> t(rowlen.size());
	for (size_t r = 0; r < rowlen.size(); r++) t[r] = (T)rowlen[r];
	fwrite(t.data(), sizeof(T), t.size(), f);
}

template <typename T>
void write_cols(FILE* f, const std::vector<uint32_t>& ja)
{
	std::vector<T> t(ja.begin(), ja.end());
	fwrite(t.data(), sizeof(T), t.size(), f);
}

void write_csr_tail(FILE* f, int ncols, const std::vector<int64_t>& rowlen, const std::vector<uint32_t>& ja)
{
	int w = width_for(ncols);
	if (w == 1) { write_deltas<uint8_t>(f, rowlen); write_cols<uint8_t>(f, ja); }
	else if (w == 2) { write_deltas<uint16_t>(f, rowlen); write_cols<uint16_t>(f, ja); }
	else { write_deltas<uint32_t>(f, rowlen); write_cols<uint32_t>(f, ja); }
}

template <typename T>
T* alloc_array(size_t n)
{
	return (T*)calloc(n ? n : 1, sizeof(T));
}

} // namespace

extern "C" void hddm_lb_free(void* p) { free(p); }

extern "C" int hddm_lb_plt_read(const char* path, int nagents, int* dim_out, int* nno_out, int* dofs_out,
	int* level_out, int32_t** index_i, int32_t** index_j, double** surplus)
{
	if (!path || !dim_out || !nno_out || !dofs_out || !level_out || !index_i || !index_j || !surplus)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plt_read: NULL argument");
	*index_i = nullptr; *index_j = nullptr; *surplus = nullptr;

	File bin(fopen(path, "rb"));
	if (!bin.f) return set_error(HDDM_LB_ERROR_IO, "Error opening file: %s", path);
	char marker[10];
	if (!read_exact(bin.f, marker, sizeof(marker)))
		return set_error(HDDM_LB_ERROR_IO, "Error reading file: %s", path);
	const bool compressed = memcmp(marker, "compressed", 10) == 0;
	if (!compressed) rewind(bin.f);

	int32_t hdr[4];
	if (!read_exact(bin.f, hdr, sizeof(hdr)))
		return set_error(HDDM_LB_ERROR_IO, "Error reading file: %s", path);
	int dim = hdr[0], nno = hdr[1], dofs = hdr[2], level = hdr[3];

	std::unique_ptr<File> text;
	bool binaryio = true;
	if (!compressed && dim != nagents)
	{
		// Guess text format, as the reference does when dim mismatches nagents.
		text.reset(new File(fopen(path, "r")));
		if (!text->f) return set_error(HDDM_LB_ERROR_IO, "Error opening file: %s", path);
		binaryio = false;
		if (fscanf(text->f, "%d %d %d %d", &dim, &nno, &dofs, &level) != 4)
			return set_error(HDDM_LB_ERROR_IO, "Error reading file: %s", path);
	}
	if (dim != nagents)
		return set_error(HDDM_LB_ERROR_FORMAT, "File \"%s\" # of dimensions (%d) mismatches config (%d)", path, dim, nagents);
	if (nno < 0 || dofs < 0 || dim < 1)
		return set_error(HDDM_LB_ERROR_FORMAT, "%s: bad header dim=%d nno=%d dofs=%d", path, dim, nno, dofs);

	int32_t* ii = alloc_array<int32_t>((size_t)nno * dim);
	int32_t* jj = alloc_array<int32_t>((size_t)nno * dim);
	double* ss = alloc_array<double>((size_t)nno * dofs);
	auto fail = [&](int code) { free(ii); free(jj); free(ss); return code; };
	if (!ii || !jj || !ss) return fail(set_error(HDDM_LB_ERROR_OUT_OF_MEMORY, "plt_read: out of memory"));

	if (compressed)
	{
		FILE* f = bin.f;
		uint32_t nnz = 0;
		if (!read_marker(f, "index") || !read_exact(f, &nnz, sizeof(nnz)))
			return fail(set_error(HDDM_LB_ERROR_FORMAT, "%s: Cannot read index data", path));
		std::vector<uint16_t> pairs((size_t)nnz * 2);
		std::vector<int64_t> ia; std::vector<uint32_t> ja;
		if (!read_exact(f, pairs.data(), sizeof(uint16_t) * pairs.size()) ||
			!read_csr_tail(f, nno, dim, nnz, ia, ja))
			return fail(set_error(HDDM_LB_ERROR_FORMAT, "%s: truncated index data", path));
		int rc = check_csr(path, "index", nno, dim, nnz, ia, ja);
		if (rc) return fail(rc);
		for (int r = 0; r < nno; r++)
			for (int64_t e = ia[r]; e < ia[r + 1]; e++)
			{
				ii[(int64_t)r * dim + ja[e]] = pairs[2 * e];
				jj[(int64_t)r * dim + ja[e]] = pairs[2 * e + 1];
			}

		uint32_t snnz = 0;
		if (!read_marker(f, "surplus") || !read_exact(f, &snnz, sizeof(snnz)))
			return fail(set_error(HDDM_LB_ERROR_FORMAT, "%s: Cannot read surplus data", path));
		std::vector<double> vals(snnz);
		if (!read_exact(f, vals.data(), sizeof(double) * snnz) ||
			!read_csr_tail(f, nno, dofs, snnz, ia, ja))
			return fail(set_error(HDDM_LB_ERROR_FORMAT, "%s: truncated surplus data", path));
		rc = check_csr(path, "surplus", nno, dofs, snnz, ia, ja);
		if (rc) return fail(rc);
		for (int r = 0; r < nno; r++)
			for (int64_t e = ia[r]; e < ia[r + 1]; e++)
				ss[(int64_t)r * dofs + ja[e]] = vals[e];
	}
	else
	{
		std::vector<int32_t> row(2 * (size_t)dim);
		std::vector<double> srow(dofs);
		for (int r = 0; r < nno; r++)
		{
			if (binaryio)
			{
				if (!read_exact(bin.f, row.data(), sizeof(int32_t) * row.size()) ||
					!read_exact(bin.f, srow.data(), sizeof(double) * srow.size()))
					return fail(set_error(HDDM_LB_ERROR_FORMAT, "Error reading file: %s", path));
			}
			else
			{
				for (size_t e = 0; e < row.size(); e++)
					if (fscanf(text->f, "%d", &row[e]) != 1)
						return fail(set_error(HDDM_LB_ERROR_FORMAT, "Error reading file: %s", path));
				for (int c = 0; c < dofs; c++)
					if (fscanf(text->f, "%lf", &srow[c]) != 1)
						return fail(set_error(HDDM_LB_ERROR_FORMAT, "Error reading file: %s", path));
			}
			for (int d = 0; d < dim; d++)
			{
				int l = row[d];
				if (l < 1 || l > 31)
					return fail(set_error(HDDM_LB_ERROR_FORMAT, "%s: node %d dim %d has level %d", path, r, d, l));
				int i = (l == 1) ? 0 : (1 << (l - 1));
				int j = row[dim + d] - 1;
				if (i == 0) j = 0;
				ii[(int64_t)r * dim + d] = i;
				jj[(int64_t)r * dim + d] = j;
			}
			memcpy(ss + (int64_t)r * dofs, srow.data(), sizeof(double) * dofs);
		}
	}

	*dim_out = dim; *nno_out = nno; *dofs_out = dofs; *level_out = level;
	*index_i = ii; *index_j = jj; *surplus = ss;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_plt_write_compressed(const char* path, int dim, int nno, int dofs, int level,
	const int32_t* index_i, const int32_t* index_j, const double* surplus)
{
	if (!path || dim < 1 || nno < 0 || dofs < 0 || (nno && (!index_i || !index_j || !surplus)))
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plt_write_compressed: bad arguments");
	File out(fopen(path, "wb"));
	if (!out.f) return set_error(HDDM_LB_ERROR_IO, "Error opening file for writing: %s", path);
	FILE* f = out.f;
	fwrite("compressed", 1, 10, f);
	int32_t hdr[4] = { dim, nno, dofs, level };
	fwrite(hdr, sizeof(int32_t), 4, f);

	std::vector<uint16_t> pairs;
	std::vector<int64_t> rowlen(nno + 1, 0);
	std::vector<uint32_t> ja;
	for (int r = 0; r < nno; r++)
		for (int d = 0; d < dim; d++)
		{
			int32_t i = index_i[(int64_t)r * dim + d], j = index_j[(int64_t)r * dim + d];
			if (!i && !j) continue;
			if (i < 0 || j < 0 || i > 65535 || j > 65535)
				return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plt_write_compressed: (i, j) = (%d, %d) does not fit u16", i, j);
			pairs.push_back((uint16_t)i); pairs.push_back((uint16_t)j);
			ja.push_back((uint32_t)d);
			rowlen[r + 1]++;
		}
	fwrite("index", 1, 5, f);
	uint32_t nnz = (uint32_t)ja.size();
	fwrite(&nnz, sizeof(nnz), 1, f);
	fwrite(pairs.data(), sizeof(uint16_t), pairs.size(), f);
	write_csr_tail(f, dim, rowlen, ja);

	std::vector<double> vals;
	std::fill(rowlen.begin(), rowlen.end(), 0);
	ja.clear();
	for (int r = 0; r < nno; r++)
		for (int c = 0; c < dofs; c++)
		{
			double v = surplus[(int64_t)r * dofs + c];
			if (v == 0.0) continue;
			vals.push_back(v); ja.push_back((uint32_t)c); rowlen[r + 1]++;
		}
	fwrite("surplus", 1, 7, f);
	uint32_t snnz = (uint32_t)vals.size();
	fwrite(&snnz, sizeof(snnz), 1, f);
	fwrite(vals.data(), sizeof(double), vals.size(), f);
	write_csr_tail(f, dofs, rowlen, ja);
	if (ferror(f)) return set_error(HDDM_LB_ERROR_IO, "Error writing file: %s", path);
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_plt_write_uncompressed(const char* path, int binary, int dim, int nno, int dofs,
	int level, const int32_t* levels, const int32_t* indices, const double* surplus)
{
	if (!path || dim < 1 || nno < 0 || dofs < 0 || (nno && (!levels || !indices || !surplus)))
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "plt_write_uncompressed: bad arguments");
	File out(fopen(path, binary ? "wb" : "w"));
	if (!out.f) return set_error(HDDM_LB_ERROR_IO, "Error opening file for writing: %s", path);
	FILE* f = out.f;
	int32_t hdr[4] = { dim, nno, dofs, level };
	if (binary) fwrite(hdr, sizeof(int32_t), 4, f);
	else fprintf(f, "%d %d %d %d\n", dim, nno, dofs, level);
	for (int r = 0; r < nno; r++)
	{
		const int32_t* lr = levels + (int64_t)r * dim;
		const int32_t* xr = indices + (int64_t)r * dim;
		const double* sr = surplus + (int64_t)r * dofs;
		if (binary)
		{
			fwrite(lr, sizeof(int32_t), dim, f);
			fwrite(xr, sizeof(int32_t), dim, f);
			fwrite(sr, sizeof(double), dofs, f);
		}
		else
		{
			for (int d = 0; d < dim; d++) fprintf(f, "%d ", lr[d]);
			for (int d = 0; d < dim; d++) fprintf(f, "%d ", xr[d]);
			for (int c = 0; c < dofs; c++) fprintf(f, "%.17g ", sr[c]);
			fprintf(f, "\n");
		}
	}
	if (ferror(f)) return set_error(HDDM_LB_ERROR_IO, "Error writing file: %s", path);
	return HDDM_LB_SUCCESS;
}
