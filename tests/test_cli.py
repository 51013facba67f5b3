"""The command-line tool and the .symb format on the CPU: option handling that
ends before any device work, and the file layout of include/esl_symb.hpp
against an independent reader/writer of the reference's layout."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from . import symb_file
from .helpers import ROOT

CLI = os.path.join(ROOT, "paraexastolog_amd", "exastolog")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(CLI), reason="exastolog is not built")
@pytest.mark.parametrize("args,msg", [
    (["--bogus"], "Unknown option"),
    ([], "Usage"),
    (["-f", "x.bnet", "-t", "GAUSS"], "Bad transition distribution type"),
    (["-f", "x.bnet", "-i", "RANDOM"], "RANDOM initial state distribution not implemented"),
    (["-f", "x.bnet", "-i", "OTHER"], "Bad initial state distribution type"),
    (["-f", "x.bnet", "--initial-fixed-prob", "1.5"], "Bad fixed initial states probability"),
    (["-f", "x.bnet", "--trans-rates", "x_A=1"], "Bad --trans-rates"),
    (["-f", "x.bnet", "--trans-rates", "u_A=abc"], "Bad --trans-rates"),
    (["-f", "x.bnet", "--initial-fixed-list", "A=2"], "Bad --initial-fixed-list"),
    (["-f"], "Missing value"),
])
def test_cli_rejects_bad_options(args, msg):
    # main.cpp:214-263: every such error prints a reason and exits with 1
    r = run(*args)
    assert r.returncode == 1
    assert msg in r.stderr


HARNESS = r"""
#include <cstdio>
#include <iostream>
#include "esl_symb.hpp"
int main(int argc, char** argv) {
  std::string mode = argv[1], path = argv[2];
  if (mode == "write") {
    esl_symb::data_t d;
    d.rows = {1, 2, 0}; d.cols = {0, 0, 1}; d.indptr = {0, 2, 3};
    d.ordered_vertices = {1, 0}; d.terminals_offsets = {0, 1}; d.nonterminals_offsets = {1, 2};
    d.rates = {1.0f, 0.0f, 2.5f, 1.0f};
    d.term_indptr = {0, 1}; d.term_indices = {1}; d.term_data = {1.0f};
    d.nonterm_indptr = {0, 2}; d.nonterm_indices = {1, 0}; d.nonterm_data = {1.0f, 1.0f};
    esl_symb::write(path, d);
    return 0;
  }
  try {
    auto d = esl_symb::read(path, mode == "read_inv");
    std::cout << d.rows.size() << " " << d.inv_data.size() << " " << d.rates[2] << " "
              << esl_symb::compatible_zero_rates(d, {3.0, 0.0, 1.0, 1.0}) << " "
              << esl_symb::compatible_zero_rates(d, {3.0, 1.0, 1.0, 1.0}) << " "
              << esl_symb::same_rates(d, {1.0, 0.0, 2.5, 1.0}) << " "
              << esl_symb::final_from_solution(d, {0.25, 0.75})[1] << std::endl;
  } catch (std::exception& e) { std::cout << "error: " << e.what() << std::endl; }
  return 0;
}
"""


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    d = tmp_path_factory.mktemp("symb")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    exe = d / "h"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    return exe


def test_symb_writer_matches_reference_layout(harness, tmp_path):
    p = tmp_path / "a.symb"
    subprocess.run([str(harness), "write", str(p)], check=True)
    d = symb_file.read(p)  # independent reader: counts, order, types and total length
    assert d["indptr"].tolist() == [0, 2, 3] and d["ordered_vertices"].tolist() == [1, 0]
    assert d["rates"].dtype == np.float32 and d["rates"].tolist() == [1.0, 0.0, 2.5, 1.0]
    assert d["nonterm_indices"].tolist() == [1, 0] and d["inv_data"].size == 0
    assert os.path.getsize(p) == 128 + 4 * (3 + 3 + 3 + 2 + 2 + 2 + 4 + 2 + 1 + 1 + 2 + 2 + 2)


def test_symb_reader_semantics(harness, tmp_path):
    p = tmp_path / "b.symb"
    subprocess.run([str(harness), "write", str(p)], check=True)
    out = subprocess.run([str(harness), "read", str(p)], capture_output=True, text=True).stdout.split()
    # zero pattern of (3,0,1,1) matches (1,0,2.5,1); (3,1,1,1) does not; final = R (L x0)
    assert out == ["3", "0", "2.5", "1", "0", "1", "1"]


def test_symb_reader_skips_or_loads_a_stored_inverse(harness, tmp_path):
    p = tmp_path / "c.symb"
    d = {"rows": [1], "cols": [0], "indptr": [0, 1, 1], "ordered_vertices": [1, 0], "terminals_offsets": [0, 1],
         "nonterminals_offsets": [1, 2], "rates": [1, 0, 1, 1], "term_indptr": [0, 1], "term_indices": [1],
         "term_data": [1], "nonterm_indptr": [0, 2], "nonterm_indices": [1, 0], "nonterm_data": [1, 1],
         "inv_indptr": [0, 1, 2], "inv_indices": [0, 1], "inv_data": [2.0, 3.0]}
    symb_file.write(p, d)
    assert subprocess.run([str(harness), "read", str(p)], capture_output=True, text=True).stdout.split()[:2] == \
        ["1", "0"]
    assert subprocess.run([str(harness), "read_inv", str(p)], capture_output=True, text=True).stdout.split()[:2] == \
        ["1", "2"]
    # a file shorter or longer than its counts is rejected (persistent_solution.cu:117-118, 148-157)
    raw = open(p, "rb").read()
    for bad in (raw[:-4], raw + b"\0\0\0\0", raw[:100]):
        open(p, "wb").write(bad)
        out = subprocess.run([str(harness), "read", str(p)], capture_output=True, text=True).stdout
        assert out.startswith("error: persistent data file is corrupted")
