"""C ABI on the CPU: the library loads, exports every entry point of
include/esl.h, and its host-side stages (model reader, rates, initial state)
agree with the oracle.  No GPU call is made here."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from oracle import pipeline as P
from oracle.bnet import read_bnet, read_bnet_text
from paraexastolog_amd import EslError, initial_state, model_builder, transition_rates
from paraexastolog_amd._lib import LIB_PATH, SIGNATURES, lib

from .helpers import MODELS, REAL_MODELS, ROOT, SMALL_DB, golden, model_path

HEADERS = [os.path.join(ROOT, "include", h) for h in ("esl.h", "esl_mlp.h")]


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(esl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB_PATH), "libesl.so is not built"
    L = C.CDLL(LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/esl.h but not exported"
    # the ctypes signature table covers every declaration
    assert sorted(n for n, _, _ in SIGNATURES) == syms


def test_version():
    assert b"gfx950" in lib().esl_version()


@pytest.mark.parametrize("name", REAL_MODELS + ["database-small/" + f for f in SMALL_DB])
def test_model_truth_tables_match_oracle(name):
    om = read_bnet(model_path(name))
    pm = model_builder().construct_model(model_path(name))
    assert list(pm.nodes) == om.nodes
    for i in range(len(om.nodes)):
        ins, tab = pm.rule(i)
        assert ins == om.inputs[i]
        assert tab.tolist() == om.tables[i].tolist()


def test_model_errors():
    b = model_builder()
    with pytest.raises(EslError) as e:
        b.construct_model("/nonexistent/model.bnet")
    assert e.value.status == 1
    for bad in ["targets, factors\nA A\n", "targets, factors\nA, (A\n", "targets, factors\nA, A &\n",
                "targets, factors\nA, \n"]:
        with pytest.raises(EslError) as e:
            b.parse_model(bad)
        assert e.value.status == 2
    big = "targets, factors\n" + "".join(f"x{i}, x{(i + 1) % 30}\n" for i in range(30))
    with pytest.raises(EslError) as e:
        b.parse_model(big)
    assert e.value.status == 3


def test_last_line_without_newline_is_kept():
    m = model_builder().parse_model("targets, factors\nA, B\nB, A")
    assert list(m.nodes) == ["A", "B"]


def test_unknown_identifier_literals_read_true():
    # the reference drops unknown identifiers from DNF clauses (model.cpp:61-70)
    m = model_builder().parse_model("targets, factors\nA, B & !Z\nB, !Y\n")
    ins, tab = m.rule(0)
    assert ins == [1] and tab.tolist() == [False, True]
    ins, tab = m.rule(1)
    assert ins == [] and tab.tolist() == [True]


def test_rates_uniform_and_overrides():
    m = model_builder().construct_model(model_path("toy.bnet"))
    r = transition_rates(m)
    r.generate_uniform()
    assert r.rates.tolist() == [1.0] * 6
    r.generate_uniform([("A", 3.0)], [("C", 2.0)])
    assert r.rates.tolist() == [3.0, 1.0, 1.0, 2.0, 1.0, 1.0]
    with pytest.raises(EslError):
        r.generate_uniform([("Q", 1.0)])
    with pytest.raises(EslError):
        r.generate_uniform([], [("A", -1.0)])
    om = read_bnet(model_path("toy.bnet"))
    assert P.rates_uniform(om, {"A": 3.0}, {"C": 2.0}).tolist() == r.rates.tolist()


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_normal_rates_bit_compatible_with_libstdcxx(tmp_path):
    # transition_rates::generate_normal (transition_rates.cpp:75-97): default_random_engine
    # + normal_distribution<float>, negatives redrawn; same libstdc++ -> same stream
    src = tmp_path / "n.cpp"
    src.write_text(r"""
#include <cstdio>
#include <random>
int main(){ std::default_random_engine g; std::normal_distribution<float> d(1.0f, 0.7f);
 for(int i=0;i<6;i++){ float x; int c=0; do { x=d(g); c++; } while(x<0 && c<1000); std::printf("%.9g\n",(double)x);} }
""")
    exe = tmp_path / "n"
    subprocess.run(["g++", "-O1", str(src), "-o", str(exe)], check=True)
    ref = [float(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    m = model_builder().construct_model(model_path("toy.bnet"))
    r = transition_rates(m)
    r.generate_normal(1.0, 0.7)
    np.testing.assert_array_equal(r.rates.astype(np.float32), np.array(ref, dtype=np.float32))


@pytest.mark.parametrize("case", ["initial_value.toy", "solver.kras", "solver.mammal", "solver.toy2"])
def test_initial_state_matches_oracle(case):
    c = golden()[case]
    om = read_bnet(model_path(c["model"]))
    pm = model_builder().construct_model(model_path(c["model"]))
    a = P.initial_state(om, c["fixed"], c["vals"], c["p"])
    b = initial_state(pm.nodes, c["fixed"], [bool(v) for v in c["vals"]], c["p"]).state
    np.testing.assert_array_equal(a, b)


def test_initial_state_partial_probability_and_errors():
    pm = model_builder().construct_model(model_path("toy.bnet"))
    om = read_bnet(model_path("toy.bnet"))
    a = P.initial_state(om, ["A"], [1], 0.7)
    b = initial_state(pm.nodes, ["A"], [True], 0.7).state
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-16)
    assert abs(b.sum() - 1.0) < 1e-15
    np.testing.assert_array_equal(initial_state(pm.nodes).state, np.full(8, 1 / 8))
    with pytest.raises(EslError):
        initial_state(pm.nodes, ["Q"], [True])
    with pytest.raises(EslError):
        initial_state(pm.nodes, ["A", "A"], [True, False])
    with pytest.raises(EslError):
        initial_state(pm.nodes, ["A"], [True, False])


_names = ["a", "b", "c", "d"]


def _formula():
    leaf = st.sampled_from(_names)
    return st.recursive(
        leaf,
        lambda ch: st.one_of(
            st.builds(lambda x: "!" + x, ch),
            st.builds(lambda x: "(" + x + ")", ch),
            st.builds(lambda x, y: x + " & " + y, ch, ch),
            st.builds(lambda x, y: x + "|" + y, ch, ch),
        ),
        max_leaves=8,
    )


@settings(max_examples=150, deadline=None)
@given(_formula())
def test_formula_fuzz_matches_oracle(f):
    text = "targets, factors\n" + "".join(f"{v}, {f}\n" for v in _names)
    om = read_bnet_text(text)
    pm = model_builder().parse_model(text)
    ins, tab = pm.rule(0)
    assert ins == om.inputs[0]
    assert tab.tolist() == om.tables[0].tolist()
