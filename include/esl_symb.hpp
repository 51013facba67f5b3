/*
 * esl_symb.hpp -- the reference's persistent symbolic file (".symb"), header-only.
 *
 * Same bytes as persistent_solution::serialize / deserialize
 * (cuda/src/persistent_solution.cu:31-199), so files move between the two
 * implementations: 16 size_t counts, then the arrays in this order --
 *   int32 rows, cols, indptr (the transition table), ordered_vertices,
 *         terminals_offsets (T+1), nonterminals_offsets (S-T+1)
 *   float rates (2N)
 *   int32 solution_term.indptr, .indices;  float solution_term.data
 *   int32 solution_nonterm.indptr, .indices;  float solution_nonterm.data
 *   [int32 n_inverse.indptr, .indices;  float n_inverse.data]
 * The last three counts are zero when the inverse is not stored.  A reader
 * checks the file length against the counts (persistent_solution.cu:117-118,
 * 148-157).  This library never writes an inverse (its solver factors nothing);
 * files that hold one are read and the inverse is skipped.
 */
#ifndef ESL_SYMB_HPP
#define ESL_SYMB_HPP

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace esl_symb {

struct data_t
{
	std::vector<int32_t> rows, cols, indptr;
	std::vector<int32_t> ordered_vertices, terminals_offsets, nonterminals_offsets;
	std::vector<float> rates;
	std::vector<int32_t> term_indptr, term_indices;
	std::vector<float> term_data;
	std::vector<int32_t> nonterm_indptr, nonterm_indices;
	std::vector<float> nonterm_data;
	std::vector<int32_t> inv_indptr, inv_indices; // read only when asked for
	std::vector<float> inv_data;
};

namespace detail {

struct file_t
{
	std::FILE* f = nullptr;
	file_t(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode))
	{
		if (!f)
			throw std::runtime_error("cannot open " + path);
	}
	~file_t()
	{
		if (f)
			std::fclose(f);
	}
};

template <typename T>
void put(std::FILE* f, const std::vector<T>& v)
{
	if (!v.empty() && std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size())
		throw std::runtime_error("short write of the symbolic file");
}

template <typename T>
const unsigned char* take(const unsigned char* p, std::vector<T>& v, size_t count)
{
	v.resize(count);
	if (count)
		std::memcpy(v.data(), p, count * sizeof(T));
	return p + count * sizeof(T);
}

} // namespace detail

inline void write(const std::string& path, const data_t& d)
{
	detail::file_t fh(path, "wb");
	const size_t sizes[16] = { d.rows.size(),
							   d.cols.size(),
							   d.indptr.size(),
							   d.ordered_vertices.size(),
							   d.terminals_offsets.size(),
							   d.nonterminals_offsets.size(),
							   d.rates.size(),
							   d.term_indptr.size(),
							   d.term_indices.size(),
							   d.term_data.size(),
							   d.nonterm_indptr.size(),
							   d.nonterm_indices.size(),
							   d.nonterm_data.size(),
							   0,
							   0,
							   0 };
	if (std::fwrite(sizes, sizeof(size_t), 16, fh.f) != 16)
		throw std::runtime_error("short write of the symbolic file");
	detail::put(fh.f, d.rows);
	detail::put(fh.f, d.cols);
	detail::put(fh.f, d.indptr);
	detail::put(fh.f, d.ordered_vertices);
	detail::put(fh.f, d.terminals_offsets);
	detail::put(fh.f, d.nonterminals_offsets);
	detail::put(fh.f, d.rates);
	detail::put(fh.f, d.term_indptr);
	detail::put(fh.f, d.term_indices);
	detail::put(fh.f, d.term_data);
	detail::put(fh.f, d.nonterm_indptr);
	detail::put(fh.f, d.nonterm_indices);
	detail::put(fh.f, d.nonterm_data);
}

inline data_t read(const std::string& path, bool load_inverse)
{
	std::vector<unsigned char> buf;
	{
		detail::file_t fh(path, "rb");
		unsigned char tmp[1 << 16];
		size_t got;
		while ((got = std::fread(tmp, 1, sizeof(tmp), fh.f)) > 0)
			buf.insert(buf.end(), tmp, tmp + got);
	}
	if (buf.size() < 16 * sizeof(size_t))
		throw std::runtime_error("persistent data file is corrupted");
	size_t s[16];
	std::memcpy(s, buf.data(), sizeof(s));
	// 12 int32 arrays and 4 float arrays, as in the count order above
	const size_t idx = s[0] + s[1] + s[2] + s[3] + s[4] + s[5] + s[7] + s[8] + s[10] + s[11] + s[13] + s[14];
	const size_t real = s[6] + s[9] + s[12] + s[15];
	for (size_t k = 0; k < 16; ++k)
		if (s[k] > buf.size())
			throw std::runtime_error("persistent data file is corrupted");
	if (buf.size() != idx * sizeof(int32_t) + real * sizeof(float) + 16 * sizeof(size_t))
		throw std::runtime_error("persistent data file is corrupted");
	data_t d;
	const unsigned char* p = buf.data() + 16 * sizeof(size_t);
	p = detail::take(p, d.rows, s[0]);
	p = detail::take(p, d.cols, s[1]);
	p = detail::take(p, d.indptr, s[2]);
	p = detail::take(p, d.ordered_vertices, s[3]);
	p = detail::take(p, d.terminals_offsets, s[4]);
	p = detail::take(p, d.nonterminals_offsets, s[5]);
	p = detail::take(p, d.rates, s[6]);
	p = detail::take(p, d.term_indptr, s[7]);
	p = detail::take(p, d.term_indices, s[8]);
	p = detail::take(p, d.term_data, s[9]);
	p = detail::take(p, d.nonterm_indptr, s[10]);
	p = detail::take(p, d.nonterm_indices, s[11]);
	p = detail::take(p, d.nonterm_data, s[12]);
	if (load_inverse)
	{
		p = detail::take(p, d.inv_indptr, s[13]);
		p = detail::take(p, d.inv_indices, s[14]);
		p = detail::take(p, d.inv_data, s[15]);
	}
	return d;
}

// persistent_solution::has_compatible_zero_rates (persistent_solution.cu:225-235):
// the same rates are zero (rates compared as the reference stores them, FP32)
inline bool compatible_zero_rates(const data_t& d, const std::vector<double>& rates)
{
	if (d.rates.size() != rates.size())
		throw std::runtime_error("symbolic data dimensions do not match model dimensions");
	for (size_t i = 0; i < rates.size(); ++i)
		if ((d.rates[i] == 0.0f) != ((float)rates[i] == 0.0f))
			return false;
	return true;
}

// persistent_solution::are_same (persistent_solution.cu:237-245)
inline bool same_rates(const data_t& d, const std::vector<double>& rates)
{
	if (d.rates.size() != rates.size())
		throw std::runtime_error("symbolic data dimensions do not match model dimensions");
	for (size_t i = 0; i < rates.size(); ++i)
		if (d.rates[i] != (float)rates[i])
			return false;
	return true;
}

// final_state = R (L x0) from a stored solution (solver.cu:302-309, FP64 sums of
// the stored FP32 values)
inline std::vector<double> final_from_solution(const data_t& d, const std::vector<double>& x0)
{
	const size_t T = d.term_indptr.empty() ? 0 : d.term_indptr.size() - 1;
	if (d.nonterm_indptr.size() != T + 1)
		throw std::runtime_error("persistent data file is corrupted");
	std::vector<double> y(T, 0.0), fin(x0.size(), 0.0);
	for (size_t j = 0; j < T; ++j)
		for (int32_t k = d.nonterm_indptr[j]; k < d.nonterm_indptr[j + 1]; ++k)
		{
			const int32_t c = d.nonterm_indices.at(k);
			y[j] += (double)d.nonterm_data.at(k) * x0.at(c);
		}
	for (size_t j = 0; j < T; ++j)
		for (int32_t k = d.term_indptr[j]; k < d.term_indptr[j + 1]; ++k)
			fin.at(d.term_indices.at(k)) = (double)d.term_data.at(k) * y[j];
	return fin;
}

} // namespace esl_symb

#endif /* ESL_SYMB_HPP */
