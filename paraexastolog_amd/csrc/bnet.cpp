// .bnet reader and rule compiler (host C++).
//
// Line handling follows model_builder::construct_model
// (reference cuda/src/model.cpp:80-125): an optional first line starting with
// "targets" is skipped; the target is the text before the first comma and the
// factor starts after the spaces that follow it (model.cpp:100-104).  The
// formula grammar is the one of BoolExprParser
// (cuda/src/boolstuff/BoolExprParser.cpp:77-200):
//     expr   := term ( '|' expr )?        (right associative)
//     term   := factor ( '&' term )?
//     factor := '!'* atom
//     atom   := '(' expr ')' | identifier [A-Za-z0-9_]+
// with every isspace() character skipped before a token.
//
// Instead of a DNF (BoolExpr.tpp:479-555, clause masks in model.cpp:35-78) each
// rule is compiled to a truth table over the nodes it reads; the table is what
// the device kernel evaluates.  For formulas over known node names the two are
// the same Boolean function.  The reference drops identifiers that are not node
// names from their DNF clause (model.cpp:61-70); here each literal of such an
// identifier, after negations are pushed to the leaves, reads as true, which is
// the same thing clause by clause.
//
// On purpose unlike the reference: a last line without '\n' is kept (the
// reference loses it, model.cpp:98-110) and malformed lines are reported as
// ESL_E_PARSE instead of escaping as std::out_of_range.

#include <algorithm>
#include <cctype>
#include <map>
#include <memory>

#include "model.hpp"

namespace esl {
namespace {

enum class op_t
{
	var,
	konst,
	not_,
	and_,
	or_
};

struct node_t
{
	op_t op;
	int var = -1; // node index, -1 = unknown identifier
	bool value = false;
	std::unique_ptr<node_t> a, b;
};
using node_p = std::unique_ptr<node_t>;

class parser_t
{
	const std::string& s_;
	const std::map<std::string, int>& names_;
	size_t i_ = 0;

	void skip()
	{
		while (i_ < s_.size() && std::isspace((unsigned char)s_[i_]))
			++i_;
	}
	bool seen(char c)
	{
		skip();
		return i_ < s_.size() && s_[i_] == c;
	}
	[[noreturn]] void fail(const std::string& what)
	{
		throw error(2, "formula '" + s_ + "': " + what + " at offset " + std::to_string(i_));
	}

	node_p expr()
	{
		node_p l = term();
		if (!seen('|'))
			return l;
		++i_;
		auto n = std::make_unique<node_t>();
		n->op = op_t::or_;
		n->a = std::move(l);
		n->b = expr();
		return n;
	}
	node_p term()
	{
		node_p l = factor();
		if (!seen('&'))
			return l;
		++i_;
		auto n = std::make_unique<node_t>();
		n->op = op_t::and_;
		n->a = std::move(l);
		n->b = term();
		return n;
	}
	node_p factor()
	{
		bool neg = false;
		while (seen('!'))
		{
			++i_;
			neg = !neg;
		}
		node_p a = atom();
		if (!neg)
			return a;
		auto n = std::make_unique<node_t>();
		n->op = op_t::not_;
		n->a = std::move(a);
		return n;
	}
	node_p atom()
	{
		skip();
		if (seen('('))
		{
			++i_;
			node_p e = expr();
			if (!seen(')'))
				fail("runaway parenthesis");
			++i_;
			return e;
		}
		size_t start = i_;
		while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_'))
			++i_;
		if (i_ == start)
			fail("identifier expected");
		auto n = std::make_unique<node_t>();
		n->op = op_t::var;
		auto it = names_.find(s_.substr(start, i_ - start));
		n->var = it == names_.end() ? -1 : it->second;
		return n;
	}

public:
	parser_t(const std::string& s, const std::map<std::string, int>& names) : s_(s), names_(names) {}

	node_p parse()
	{
		node_p e = expr();
		skip();
		if (i_ != s_.size())
			fail("garbage at end");
		return e;
	}
};

// Negation normal form; a literal of an unknown identifier becomes true.
node_p nnf(const node_t& n, bool neg)
{
	auto r = std::make_unique<node_t>();
	switch (n.op)
	{
		case op_t::var:
			if (n.var < 0)
			{
				r->op = op_t::konst;
				r->value = true;
				return r;
			}
			if (neg)
			{
				r->op = op_t::not_;
				r->a = std::make_unique<node_t>();
				r->a->op = op_t::var;
				r->a->var = n.var;
				return r;
			}
			r->op = op_t::var;
			r->var = n.var;
			return r;
		case op_t::konst:
			r->op = op_t::konst;
			r->value = n.value != neg;
			return r;
		case op_t::not_:
			return nnf(*n.a, !neg);
		case op_t::and_:
		case op_t::or_:
			r->op = (n.op == op_t::and_) != neg ? op_t::and_ : op_t::or_;
			r->a = nnf(*n.a, neg);
			r->b = nnf(*n.b, neg);
			return r;
	}
	return r;
}

void collect(const node_t& n, std::vector<int32_t>& vars)
{
	if (n.op == op_t::var)
		vars.push_back(n.var);
	if (n.a)
		collect(*n.a, vars);
	if (n.b)
		collect(*n.b, vars);
}

// Evaluates the formula for 64 assignments at once (one bit per assignment).
uint64_t eval64(const node_t& n, const std::vector<uint64_t>& var_bits)
{
	switch (n.op)
	{
		case op_t::var:
			return var_bits[n.var];
		case op_t::konst:
			return n.value ? ~0ull : 0ull;
		case op_t::not_:
			return ~eval64(*n.a, var_bits);
		case op_t::and_:
			return eval64(*n.a, var_bits) & eval64(*n.b, var_bits);
		case op_t::or_:
			return eval64(*n.a, var_bits) | eval64(*n.b, var_bits);
	}
	return 0;
}

rule_t compile(const node_t& f, int n_nodes)
{
	rule_t r;
	collect(f, r.inputs);
	std::sort(r.inputs.begin(), r.inputs.end());
	r.inputs.erase(std::unique(r.inputs.begin(), r.inputs.end()), r.inputs.end());
	const int k = (int)r.inputs.size();
	if (k > max_rule_inputs)
		throw error(2, "rule reads " + std::to_string(k) + " nodes; at most " + std::to_string(max_rule_inputs) +
						   " are supported");
	const uint64_t entries = 1ull << k;
	r.table.assign(entries, 0);
	std::vector<uint64_t> var_bits(n_nodes, 0);
	for (uint64_t base = 0; base < entries; base += 64)
	{
		for (int j = 0; j < k; ++j)
		{
			uint64_t w = 0;
			for (int l = 0; l < 64; ++l)
				if (((base + l) >> j) & 1)
					w |= 1ull << l;
			var_bits[r.inputs[j]] = w;
		}
		const uint64_t v = eval64(f, var_bits);
		for (uint64_t l = 0; l < 64 && base + l < entries; ++l)
			r.table[base + l] = (v >> l) & 1;
	}
	return r;
}

} // namespace

model_t parse_bnet(const std::string& text)
{
	std::vector<std::string> lines;
	size_t pos = 0;
	while (pos < text.size())
	{
		size_t nl = text.find('\n', pos);
		if (nl == std::string::npos)
			nl = text.size();
		lines.push_back(text.substr(pos, nl - pos));
		pos = nl + 1;
	}
	size_t first = 0;
	if (!lines.empty() && lines[0].rfind("targets", 0) == 0)
		first = 1;

	model_t m;
	for (size_t li = first; li < lines.size(); ++li)
	{
		const std::string& ln = lines[li];
		size_t c = ln.find(',');
		if (c == std::string::npos)
			throw error(2, "line " + std::to_string(li + 1) + " has no comma: '" + ln + "'");
		size_t fs = ln.find_first_not_of(' ', c + 1);
		if (fs == std::string::npos)
			throw error(2, "line " + std::to_string(li + 1) + " has an empty factor");
		m.nodes.push_back(ln.substr(0, c));
		m.formulas.push_back(ln.substr(fs));
	}
	if (m.nodes.empty())
		throw error(2, "model has no nodes");
	if ((int)m.nodes.size() > max_nodes)
		throw error(3, "model has " + std::to_string(m.nodes.size()) + " nodes; at most " +
						   std::to_string(max_nodes) + " are supported (2^N states)");

	std::map<std::string, int> names;
	for (int i = 0; i < (int)m.nodes.size(); ++i)
		names.emplace(m.nodes[i], i); // first occurrence wins, as std::find in the reference

	for (const auto& f : m.formulas)
	{
		node_p ast = parser_t(f, names).parse();
		node_p norm = nnf(*ast, false);
		m.rules.push_back(compile(*norm, (int)m.nodes.size()));
	}
	return m;
}

} // namespace esl
