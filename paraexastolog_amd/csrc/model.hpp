// Host-side model: node names and one truth table per rule.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace esl {

struct error : std::runtime_error
{
	int code;
	error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// One rule f_i: the ascending node indices it reads and its truth table,
// indexed by the input bits packed in that order (input k -> bit k).
struct rule_t
{
	std::vector<int32_t> inputs;
	std::vector<uint8_t> table;
};

struct model_t
{
	std::vector<std::string> nodes;
	std::vector<std::string> formulas;
	std::vector<rule_t> rules;
};

// Reads a .bnet text: optional "targets..." header line, then one
// "<target>,<factor>" line per node.  Throws esl::error.
model_t parse_bnet(const std::string& text);

// Largest rule arity accepted (truth tables of 2^k entries).
constexpr int max_rule_inputs = 24;
// Largest model: states are 32-bit integers and the table CSR uses int32.
constexpr int max_nodes = 26;

} // namespace esl
