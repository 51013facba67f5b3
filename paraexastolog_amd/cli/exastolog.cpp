// exastolog -- command-line drop-in for the reference's cudaExaStoLog
// (cuda/src/main.cpp:165-312): same options, same output text, same .symb files.
//
//   -f, --bnet-file BNET_FILE            model (required)
//   -t, --trans-distribution T_DISTR     UNIFORM (default) or NORMAL
//       --mean MEAN, --std STD           of the NORMAL rates (defaults 0 and 1)
//       --trans-rates TRANS_RATE         comma-separated (u|d)_NAME=VAL
//   -i, --initial-distribution DISTR     UNIFORM (default); RANDOM is not implemented (as in the reference)
//       --initial-fixed-prob INIT_PROB   probability of the fixed initial states (default 1)
//       --initial-fixed-list LIST        comma-separated NAME=(0|1)
//   -s, --symbolic                       reuse / create BNET_FILE-without-extension.symb
//   -n, --no-symb-rates                  do not load a stored inverse (this library stores none)
//
// One deliberate difference: --initial-fixed-prob takes effect.  The reference
// reads it into the NORMAL standard deviation instead (main.cpp:254-256), so
// there it never changes the initial state.
#include <cctype>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <map>
#include <stdexcept>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "esl.h"
#include "esl_symb.hpp"

namespace {

const char* usage_text =
	"Usage: exastolog -f BNET_FILE [options]\n"
	"  -f, --bnet-file BNET_FILE           Path to the bnet file.\n"
	"  -t, --trans-distribution T_DISTR    Distribution of transition rates. Can be NORMAL or UNIFORM (default).\n"
	"      --mean MEAN                     Mean value for normal transition rate distribution.\n"
	"      --std STD                       Standard deviation value for normal transition rate distribution.\n"
	"      --trans-rates TRANS_RATE        Explicit transition rates. A comma separated list of (u|d)_NAME=VAL.\n"
	"  -i, --initial-distribution DISTR    Distribution of initial state. Can be RANDOM or UNIFORM (default).\n"
	"      --initial-fixed-prob INIT_PROB  Probability of fixed initial states.\n"
	"      --initial-fixed-list LIST       Specify fixed parts of the initial state. A comma separated list of "
	"NAME=(0|1).\n"
	"  -s, --symbolic                      Turns on symbolic computation (BNET_FILE.symb is created and reused).\n"
	"  -n, --no-symb-rates                 Do not load a stored inverse of the non-terminal part.\n";

struct args_t
{
	std::string bnet;
	bool uniform = true;
	float mean = 0.f, sd = 1.f, fixed_prob = 1.f;
	std::vector<std::string> up_names, dn_names, fixed_nodes;
	std::vector<double> up_vals, dn_vals;
	std::vector<uint8_t> fixed_vals;
	bool symbolic = false, no_inverse = false;
};

std::vector<std::string> split_commas(const std::string& v)
{
	std::vector<std::string> out;
	size_t a = 0;
	for (;;)
	{
		const size_t e = v.find(',', a);
		out.push_back(v.substr(a, e == std::string::npos ? std::string::npos : e - a));
		if (e == std::string::npos)
			return out;
		a = e + 1;
	}
}

// (u|d)_NAME=VAL, VAL read as a float (main.cpp:18-62)
bool parse_rates(const std::string& v, args_t& a)
{
	for (const std::string& e : split_commas(v))
	{
		const size_t eq = e.find('=');
		if (eq == std::string::npos || e.size() < 2 || (std::toupper(e[0]) != 'D' && std::toupper(e[0]) != 'U') ||
			e[1] != '_')
			return false;
		float val;
		try
		{
			val = std::stof(e.substr(eq + 1));
		}
		catch (...)
		{
			return false;
		}
		const std::string node = e.substr(2, eq - 2);
		if (std::toupper(e[0]) == 'D')
		{
			a.dn_names.push_back(node);
			a.dn_vals.push_back(val);
		}
		else
		{
			a.up_names.push_back(node);
			a.up_vals.push_back(val);
		}
	}
	return true;
}

// NAME=(0|1) (main.cpp:70-100)
bool parse_fixed(const std::string& v, args_t& a)
{
	for (const std::string& e : split_commas(v))
	{
		const size_t eq = e.find('=');
		if (eq == std::string::npos || eq + 1 >= e.size() || (e[eq + 1] != '0' && e[eq + 1] != '1'))
			return false;
		a.fixed_nodes.push_back(e.substr(0, eq));
		a.fixed_vals.push_back(e[eq + 1] == '1');
	}
	return true;
}

bool parse_float(const std::string& s, float& out)
{
	try
	{
		size_t used = 0;
		out = std::stof(s, &used);
		return used == s.size();
	}
	catch (...)
	{
		return false;
	}
}

// returns 0 on success, 1 after printing the reason and the help
int parse_args(int argc, char** argv, args_t& a)
{
	const std::map<std::string, std::string> names = {
		{ "f", "bnet-file" },		  { "bnet-file", "bnet-file" },
		{ "t", "trans-distribution" }, { "trans-distribution", "trans-distribution" },
		{ "mean", "mean" },			  { "std", "std" },
		{ "trans-rates", "trans-rates" }, { "i", "initial-distribution" },
		{ "initial-distribution", "initial-distribution" }, { "initial-fixed-prob", "initial-fixed-prob" },
		{ "initial-fixed-list", "initial-fixed-list" }, { "s", "symbolic" },
		{ "symbolic", "symbolic" },	  { "n", "no-symb-rates" },
		{ "no-symb-rates", "no-symb-rates" }
	};
	std::map<std::string, std::string> got;
	for (int k = 1; k < argc; ++k)
	{
		std::string arg = argv[k], key, val;
		bool has_val = false;
		if (arg.rfind("--", 0) == 0)
			key = arg.substr(2);
		else if (arg.size() >= 2 && arg[0] == '-')
			key = arg.substr(1);
		else
		{
			std::cerr << "Unexpected argument: " << arg << "\n" << usage_text;
			return 1;
		}
		const size_t eq = key.find('=');
		if (eq != std::string::npos)
		{
			val = key.substr(eq + 1);
			key = key.substr(0, eq);
			has_val = true;
		}
		const auto it = names.find(key);
		if (it == names.end())
		{
			std::cerr << "Unknown option: " << arg << "\n" << usage_text;
			return 1;
		}
		const std::string& name = it->second;
		const bool is_flag = name == "symbolic" || name == "no-symb-rates";
		if (!is_flag && !has_val)
		{
			if (k + 1 >= argc)
			{
				std::cerr << "Missing value for " << arg << "\n" << usage_text;
				return 1;
			}
			val = argv[++k];
		}
		got[name] = val;
	}
	if (!got.count("bnet-file"))
	{
		std::cerr << usage_text;
		return 1;
	}
	a.bnet = got["bnet-file"];
	if (got.count("trans-distribution"))
	{
		const std::string& d = got["trans-distribution"];
		if (d != "NORMAL" && d != "UNIFORM")
		{
			std::cerr << "Bad transition distribution type" << std::endl << usage_text;
			return 1;
		}
		a.uniform = d == "UNIFORM";
	}
	if ((got.count("mean") && !parse_float(got["mean"], a.mean)) || (got.count("std") && !parse_float(got["std"], a.sd)))
	{
		std::cerr << "Bad numeric value" << std::endl << usage_text;
		return 1;
	}
	if (got.count("trans-rates") && !parse_rates(got["trans-rates"], a))
	{
		std::cerr << "Bad --trans-rates value" << std::endl << usage_text;
		return 1;
	}
	if (got.count("initial-distribution"))
	{
		const std::string& d = got["initial-distribution"];
		if (d == "RANDOM")
		{
			std::cerr << "RANDOM initial state distribution not implemented" << std::endl;
			return 1;
		}
		if (d != "UNIFORM")
		{
			std::cerr << "Bad initial state distribution type" << std::endl << usage_text;
			return 1;
		}
	}
	if (got.count("initial-fixed-prob") && !parse_float(got["initial-fixed-prob"], a.fixed_prob))
	{
		std::cerr << "Bad numeric value" << std::endl << usage_text;
		return 1;
	}
	if (a.fixed_prob < 0.f || a.fixed_prob > 1.f)
	{
		std::cerr << "Bad fixed initial states probability" << std::endl << usage_text;
		return 1;
	}
	if (got.count("initial-fixed-list") && !parse_fixed(got["initial-fixed-list"], a))
	{
		std::cerr << "Bad --initial-fixed-list value" << std::endl << usage_text;
		return 1;
	}
	a.symbolic = got.count("symbolic") > 0;
	a.no_inverse = got.count("no-symb-rates") > 0;
	return 0;
}

struct status_error : std::runtime_error
{
	using std::runtime_error::runtime_error;
};

void ok(esl_status s, const esl_ctx* c = nullptr)
{
	if (s != ESL_OK)
		throw status_error(esl_last_error(c));
}

// solver::print_final_state (solver.cu:336-372): nonzero states as node names
// joined by " --- " ("<nil>" for state 0) and the value as the reference's
// float prints it
void print_final_state(const std::vector<double>& fin, const std::vector<std::string>& nodes)
{
	for (size_t s = 0; s < fin.size(); ++s)
	{
		if ((float)fin[s] == 0.0f)
			continue;
		bool printed = false;
		for (size_t i = 0; i < nodes.size(); ++i)
			if (s & (size_t(1) << i))
			{
				if (printed)
					std::cout << " --- ";
				std::cout << nodes[i];
				printed = true;
			}
		if (!printed)
			std::cout << "<nil>";
		std::cout << " " << (float)fin[s] << std::endl;
	}
}

std::string symb_path(const std::string& bnet)
{
	// bpp::Path::cropExtension: drop the last extension of the file name
	const size_t slash = bnet.find_last_of('/');
	const size_t dot = bnet.find_last_of('.');
	if (dot != std::string::npos && (slash == std::string::npos || dot > slash))
		return bnet.substr(0, dot) + ".symb";
	return bnet + ".symb";
}

bool exists(const std::string& p)
{
	struct stat st;
	return ::stat(p.c_str(), &st) == 0;
}

struct run_t
{
	esl_ctx* ctx = nullptr;
	esl_model* model = nullptr;
	esl_graph* g = nullptr;
	~run_t()
	{
		esl_graph_destroy(g);
		esl_model_destroy(model);
		esl_ctx_destroy(ctx);
	}
};

// everything a .symb file stores, from the graph solved with opts.full = 1
esl_symb::data_t collect(run_t& r, const std::vector<double>& rates)
{
	esl_symb::data_t d;
	int64_t sz[7];
	ok(esl_graph_query(r.g, sz));
	const int64_t n = sz[0], nnz = sz[1], S = sz[2], T = sz[3], n_t = sz[4];
	d.indptr.resize(n + 1);
	d.rows.resize(nnz);
	d.cols.resize(nnz);
	ok(esl_table_export(r.ctx, r.g, d.indptr.data(), d.rows.data(), d.cols.data()), r.ctx);
	std::vector<int32_t> offsets(S + 1);
	int32_t tc = 0;
	d.ordered_vertices.resize(n);
	ok(esl_graph_export(r.ctx, r.g, d.ordered_vertices.data(), offsets.data(), &tc, nullptr), r.ctx);
	d.terminals_offsets.assign(offsets.begin(), offsets.begin() + T + 1);
	d.nonterminals_offsets.assign(offsets.begin() + T, offsets.end());
	for (double v : rates)
		d.rates.push_back((float)v);
	int64_t nnz_l = 0;
	ok(esl_solution_export(r.ctx, r.g, nullptr, nullptr, nullptr, &nnz_l, nullptr, nullptr, nullptr), r.ctx);
	std::vector<double> td(n_t), nd(nnz_l);
	d.term_indptr.resize(T + 1);
	d.term_indices.resize(n_t);
	d.nonterm_indptr.resize(T + 1);
	d.nonterm_indices.resize(nnz_l);
	ok(esl_solution_export(r.ctx, r.g, d.term_indptr.data(), d.term_indices.data(), td.data(), &nnz_l,
						   d.nonterm_indptr.data(), d.nonterm_indices.data(), nd.data()),
	   r.ctx);
	for (double v : td)
		d.term_data.push_back((float)v);
	for (double v : nd)
		d.nonterm_data.push_back((float)v);
	return d;
}

} // namespace

int main(int argc, char** argv)
{
	args_t a;
	if (parse_args(argc, argv, a))
		return 1;
	try
	{
		run_t r;
		ok(esl_ctx_create(0, &r.ctx));
		ok(esl_model_load(a.bnet.c_str(), &r.model));
		const int N = esl_model_num_nodes(r.model);
		std::vector<std::string> nodes;
		for (int i = 0; i < N; ++i)
			nodes.push_back(esl_model_node_name(r.model, i));

		std::vector<const char*> fixed;
		for (auto& s : a.fixed_nodes)
			fixed.push_back(s.c_str());
		std::vector<double> x0(size_t(1) << N);
		ok(esl_initial_state(r.model, fixed.data(), a.fixed_vals.data(), (int)fixed.size(), a.fixed_prob, x0.data()));

		std::vector<const char*> un, dn;
		for (auto& s : a.up_names)
			un.push_back(s.c_str());
		for (auto& s : a.dn_names)
			dn.push_back(s.c_str());
		std::vector<double> rates(2 * N);
		ok(esl_rates_fill(r.model, a.uniform ? 0 : 1, a.mean, a.sd, un.data(), a.up_vals.data(), (int)un.size(),
						  dn.data(), a.dn_vals.data(), (int)dn.size(), rates.data()));

		const std::string symb = symb_path(a.bnet);
		esl_solve_opts o;
		esl_solve_opts_default(&o);
		o.full = a.symbolic ? 1 : 0; // the stored solution needs solution_term / solution_nonterm
		std::vector<double> fin(x0.size());

		bool serialize = a.symbolic;
		bool solved = false;
		if (a.symbolic && exists(symb))
		{
			esl_symb::data_t d = esl_symb::read(symb, !a.no_inverse);
			if (esl_symb::compatible_zero_rates(d, rates))
			{
				if (esl_symb::same_rates(d, rates))
				{
					// solver(ctx, persisted, ...) with unchanged rates: the stored
					// solution gives the final state (solver.cu:36-43, 302-309)
					fin = esl_symb::final_from_solution(d, x0);
					solved = true;
					serialize = false;
				}
				else
				{
					// new rates, same zero pattern: the table is rebuilt (cheap) and
					// checked against the stored one; the graph stage is imported
					ok(esl_table_build(r.ctx, r.model, rates.data(), &r.g), r.ctx);
					int64_t sz[7];
					ok(esl_graph_query(r.g, sz));
					std::vector<int32_t> ip(sz[0] + 1), rw(sz[1]);
					ok(esl_table_export(r.ctx, r.g, ip.data(), rw.data(), nullptr), r.ctx);
					if (ip == d.indptr && rw == d.rows && d.ordered_vertices.size() == (size_t)sz[0] &&
						!d.terminals_offsets.empty() && !d.nonterminals_offsets.empty())
					{
						std::vector<int32_t> offsets(d.terminals_offsets);
						offsets.insert(offsets.end(), d.nonterminals_offsets.begin() + 1, d.nonterminals_offsets.end());
						ok(esl_graph_import(r.ctx, r.g, d.ordered_vertices.data(), offsets.data(),
											(int64_t)offsets.size() - 1, (int32_t)d.terminals_offsets.size() - 1),
						   r.ctx);
						ok(esl_solve(r.ctx, r.g, rates.data(), x0.data(), fin.data(), &o, nullptr), r.ctx);
						solved = true;
					}
					else
					{
						std::cerr << "stored table of " << symb << " does not match the model; recomputing"
								  << std::endl;
						esl_graph_destroy(r.g);
						r.g = nullptr;
					}
				}
			}
		}
		if (!solved)
		{
			ok(esl_table_build(r.ctx, r.model, rates.data(), &r.g), r.ctx);
			ok(esl_graph_reorganize(r.ctx, r.g), r.ctx);
			ok(esl_solve(r.ctx, r.g, rates.data(), x0.data(), fin.data(), &o, nullptr), r.ctx);
		}
		print_final_state(fin, nodes);
		if (serialize)
			esl_symb::write(symb, collect(r, rates));
	}
	catch (std::exception& e)
	{
		std::cerr << "Program ended with an exception: " << e.what() << std::endl;
		return 1;
	}
	return 0;
}
