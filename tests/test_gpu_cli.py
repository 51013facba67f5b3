"""The command-line tool on the GPU: its printed final state against the
reference's goldens and the oracle, and the .symb flows of main.cpp:102-157
(create, reuse with the same rates, import with new rates of the same zero
pattern, recompute when the zero pattern changes)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import pipeline as P
from oracle.bnet import read_bnet

from . import symb_file
from .helpers import ROOT, golden, model_path

pytestmark = pytest.mark.gpu
CLI = os.path.join(ROOT, "paraexastolog_amd", "exastolog")


def run(*args):
    r = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def parse(out, nodes):
    """printed lines -> {state: value}"""
    res = {}
    for line in out.strip().splitlines():
        names, val = line.rsplit(" ", 1)
        s = 0 if names == "<nil>" else sum(1 << nodes.index(x) for x in names.split(" --- "))
        res[s] = float(val)
    return res


@pytest.fixture
def toy(tmp_path):
    p = tmp_path / "toy.bnet"
    shutil.copy(model_path("toy.bnet"), p)
    return str(p)


def test_cli_prints_reference_goldens(tmp_path):
    for case in ("solver.kras", "solver.toy2", "rates.toy"):
        c = golden()[case]
        m = read_bnet(model_path(c["model"]))
        args = ["-f", model_path(c["model"])]
        if c.get("fixed"):
            args += ["--initial-fixed-list", ",".join(f"{n}={v}" for n, v in zip(c["fixed"], c["vals"]))]
        rates = c.get("rates")
        if rates:
            lst = [f"u_{k}={v}" for k, v in rates["up"].items()] + [f"d_{k}={v}" for k, v in rates["down"].items()]
            if lst:
                args += ["--trans-rates", ",".join(lst)]
        got = parse(run(*args), m.nodes)
        assert sorted(got) == list(c["nonzero_indices"])
        np.testing.assert_allclose([got[s] for s in c["nonzero_indices"]], c["nonzero_data"],
                                   atol=max(c["tol"], 1e-6))


def test_cli_matches_oracle_with_normal_rates(tmp_path):
    m = read_bnet(model_path("krasmodel15vars.bnet"))
    out = run("-f", model_path("krasmodel15vars.bnet"), "-t", "NORMAL", "--mean", "1.5", "--std", "0.5")
    got = parse(out, m.nodes)
    from paraexastolog_amd import model_builder, transition_rates

    r = transition_rates(model_builder().construct_model(model_path("krasmodel15vars.bnet")))
    r.generate_normal(1.5, 0.5)
    n = 1 << len(m.nodes)
    ip, rows, _ = P.transition_table(m, r.rates)
    x = P.solve_final(n, ip, rows, r.rates, P.initial_state(m), P.transition_graph(n, ip, rows))
    dense = np.zeros(n)
    for st, v in got.items():
        dense[st] = v
    # printed as FP32 with 6 significant digits
    np.testing.assert_allclose(dense, x, rtol=1e-5, atol=1e-9)
    assert abs(dense.sum() - 1) < 1e-5


def test_cli_symbolic_flows(toy):
    nodes = read_bnet(toy).nodes
    symb = toy[:-len(".bnet")] + ".symb"
    first = parse(run("-f", toy, "-s"), nodes)
    assert os.path.exists(symb)
    d = symb_file.read(symb)
    assert d["indptr"].size == 9 and d["ordered_vertices"].size == 8 and d["rates"].tolist() == [1.0] * 6
    assert d["terminals_offsets"][-1] == d["nonterminals_offsets"][0] and d["inv_data"].size == 0
    stamp = os.path.getmtime(symb)

    # same rates, another initial state: the stored solution is reused
    # (symbolic_no_refactor.toy, test.cu:642-690)
    c = golden()["symbolic_no_refactor.toy"]
    reuse = parse(run("-f", toy, "-s", "--initial-fixed-list", "A=0,C=0,D=0"), nodes)
    assert sorted(reuse) == c["nonzero_indices"]
    np.testing.assert_allclose([reuse[s] for s in c["nonzero_indices"]], c["nonzero_data"], atol=1e-6)
    assert os.path.getmtime(symb) == stamp  # nothing to store

    # new rates, same zero pattern: graph stage imported, solution recomputed and stored
    # (symbolic_with_refactor.toy, test.cu:780-839)
    c = golden()["symbolic_with_refactor.toy"]
    ref = parse(run("-f", toy, "--trans-rates", "d_C=2"), nodes)
    new = parse(run("-f", toy, "-s", "--trans-rates", "d_C=2"), nodes)
    assert sorted(new) == c["nonzero_indices"] == sorted(ref)
    np.testing.assert_allclose([new[s] for s in c["nonzero_indices"]], c["nonzero_data"], atol=c["tol"])
    np.testing.assert_allclose([new[s] for s in sorted(ref)], [ref[s] for s in sorted(ref)], rtol=1e-6)
    assert symb_file.read(symb)["rates"].tolist() == [1.0, 1.0, 1.0, 2.0, 1.0, 1.0]

    # a rate set to zero changes the pattern: everything is recomputed and stored
    ref0 = parse(run("-f", toy, "--trans-rates", "u_C=0"), nodes)
    z = parse(run("-f", toy, "-s", "--trans-rates", "u_C=0"), nodes)
    assert z == pytest.approx(ref0)
    assert symb_file.read(symb)["rates"].tolist()[2] == 0.0
    assert first  # the first run printed a distribution
