"""Oracle (test infrastructure only): `.bnet` reader and rule evaluator.

This is the CPU restatement used to CHECK the HIP path; nothing in the product
imports it.  It restates:

* the line handling of ``model_builder::construct_model`` (cuda/src/model.cpp:80-125):
  an optional first line starting with ``targets`` is skipped, every other
  line is ``<target>,<factor>`` where the target is everything before the
  first comma and the factor starts after the spaces that follow it
  (model.cpp:100-104);
* the grammar of ``BoolExprParser`` (cuda/src/boolstuff/BoolExprParser.cpp:77-200):
  ``expr := term ('|' expr)?``, ``term := factor ('&' term)?``,
  ``factor := '!'* atom``, ``atom := '(' expr ')' | identifier``, identifiers
  are ``[A-Za-z0-9_]+`` and every ``isspace`` character is skipped;
* the clause semantics of ``construct_clauses`` (model.cpp:35-78) and
  ``evaluation_ftor`` (cuda/src/transition_table.cu:34-42): a rule is true in
  state s iff one of its DNF clauses has all positive variables set and all
  negative variables clear.  For formulas over known identifiers this is the
  truth table of the formula; an identifier that is not a node name is
  dropped from its clause (model.cpp:61-70), i.e. each of its literals reads
  as true once negations are pushed to the leaves.

Differences from the reference, on purpose: a last line without a trailing
newline is kept (model.cpp:98-110 loses it), and a malformed line raises
``BnetError`` instead of throwing ``std::out_of_range``.
"""
from __future__ import annotations

import numpy as np


class BnetError(ValueError):
    pass


# --- AST: ('var', name) | ('not', a) | ('and', a, b) | ('or', a, b) | ('const', bool)


class _Parser:
    def __init__(self, text: str):
        self.s = text
        self.i = 0

    def _skip(self):
        while self.i < len(self.s) and self.s[self.i].isspace():
            self.i += 1

    def _seen(self, tok):
        self._skip()
        return self.s.startswith(tok, self.i)

    def parse(self):
        e = self._expr()
        self._skip()
        if self.i != len(self.s):
            raise BnetError(f"garbage at end of formula at {self.i}: {self.s!r}")
        return e

    def _expr(self):
        left = self._term()
        if not self._seen("|"):
            return left
        self.i += 1
        return ("or", left, self._expr())

    def _term(self):
        left = self._factor()
        if not self._seen("&"):
            return left
        self.i += 1
        return ("and", left, self._term())

    def _factor(self):
        neg = False
        while self._seen("!"):
            self.i += 1
            neg = not neg
        a = self._atom()
        return ("not", a) if neg else a

    def _atom(self):
        self._skip()
        if self._seen("("):
            self.i += 1
            e = self._expr()
            if not self._seen(")"):
                raise BnetError(f"runaway parenthesis in {self.s!r}")
            self.i += 1
            return e
        start = self.i
        while self.i < len(self.s) and (self.s[self.i].isalnum() or self.s[self.i] == "_"):
            self.i += 1
        if self.i == start:
            raise BnetError(f"identifier expected at {start} in {self.s!r}")
        return ("var", self.s[start:self.i])


def parse_formula(text: str):
    return _Parser(text).parse()


def _nnf(ast, known, neg=False):
    """Push negations to the leaves; literals of unknown identifiers become true."""
    k = ast[0]
    if k == "var":
        if ast[1] not in known:
            return ("const", True)
        return ("not", ast) if neg else ast
    if k == "not":
        return _nnf(ast[1], known, not neg)
    if k == "const":
        return ("const", ast[1] != neg)
    op = k
    if neg:
        op = "or" if k == "and" else "and"
    return (op, _nnf(ast[1], known, neg), _nnf(ast[2], known, neg))


def _vars(ast, out):
    if ast[0] == "var":
        out.add(ast[1])
    elif ast[0] == "not":
        _vars(ast[1], out)
    elif ast[0] in ("and", "or"):
        _vars(ast[1], out)
        _vars(ast[2], out)
    return out


def _eval(ast, env):
    k = ast[0]
    if k == "var":
        return env[ast[1]]
    if k == "const":
        return np.full(next(iter(env.values())).shape if env else (1,), ast[1])
    if k == "not":
        return ~_eval(ast[1], env)
    a = _eval(ast[1], env)
    b = _eval(ast[2], env)
    return (a & b) if k == "and" else (a | b)


class Model:
    """Node names in file order plus, per node, (inputs, truth table).

    inputs[i] is the ascending list of node indices the rule of node i reads;
    table[i] is a bool array of length 2**len(inputs[i]) indexed by the input
    bits packed in that order (input k -> bit k).
    """

    def __init__(self, nodes, formulas):
        self.nodes = list(nodes)
        self.formulas = list(formulas)
        known = {n: i for i, n in enumerate(self.nodes)}
        self.inputs = []
        self.tables = []
        for f in self.formulas:
            ast = _nnf(parse_formula(f), known)
            ins = sorted(known[v] for v in _vars(ast, set()))
            K = len(ins)
            idx = np.arange(1 << K, dtype=np.int64)
            env = {self.nodes[j]: ((idx >> k) & 1).astype(bool) for k, j in enumerate(ins)}
            if K == 0:
                val = np.array([_const_value(ast)])
            else:
                val = np.asarray(_eval(ast, env), dtype=bool)
            self.inputs.append(ins)
            self.tables.append(val)

    @property
    def n_nodes(self):
        return len(self.nodes)

    def evaluate(self, i, states):
        """f_i over an int64 array of states."""
        ins = self.inputs[i]
        idx = np.zeros(states.shape, dtype=np.int64)
        for k, j in enumerate(ins):
            idx |= ((states >> j) & 1) << k
        return self.tables[i][idx]


def _const_value(ast):
    k = ast[0]
    if k == "const":
        return ast[1]
    if k == "and":
        return _const_value(ast[1]) and _const_value(ast[2])
    if k == "or":
        return _const_value(ast[1]) or _const_value(ast[2])
    if k == "not":
        return not _const_value(ast[1])
    raise BnetError("variable in constant rule")


def read_bnet_text(text: str) -> Model:
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines = lines[:-1]
    if lines and lines[0].startswith("targets"):
        lines = lines[1:]
    nodes, formulas = [], []
    for ln in lines:
        c = ln.find(",")
        if c < 0:
            raise BnetError(f"line without a comma: {ln!r}")
        target = ln[:c]
        rest = ln[c + 1:]
        factor = rest.lstrip(" ")
        if factor == "":
            raise BnetError(f"empty factor: {ln!r}")
        nodes.append(target)
        formulas.append(factor)
    if not nodes:
        raise BnetError("model without nodes")
    return Model(nodes, formulas)


def read_bnet(path: str) -> Model:
    with open(path, "r") as fh:
        return read_bnet_text(fh.read())
