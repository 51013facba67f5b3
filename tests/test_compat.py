"""The header-only C++ mirror of the reference's class API (include/esl_compat.hpp)
compiles against include/esl.h and links against libesl.so; its host-side parts
(model, rates, initial state) run without a GPU."""
import os
import shutil
import subprocess

import pytest

from paraexastolog_amd._lib import LIB_PATH

from .helpers import ROOT, model_path

PROG = r"""
#include <cstdio>
#include "esl_compat.hpp"
int main(int argc, char** argv) {
    compat::model_builder b;
    auto m = b.construct_model(argv[1]);
    compat::transition_rates r(m);
    r.generate_uniform({}, {{"C", 2.0}});
    compat::initial_state s(m, {"A", "C", "D"}, {false, false, false}, 1.0);
    std::printf("%zu %g %g %g\n", m.nodes.size(), r.rates[3], s.state[0], s.state[1]);
    if (argc > 2) {                 // device part
        compat::cu_context ctx;
        compat::transition_table t(ctx, m, r.rates);
        t.construct_table();
        compat::transition_graph g(ctx, t);
        g.reorganize_vertices();
        compat::solver sv(ctx, t, g, r, compat::initial_state(m));
        sv.solve();
        std::printf("%d %g %g %g\n", g.terminals_count, sv.final_state[1], sv.final_state[2], sv.final_state[4]);
    }
    return 0;
}
"""


def _build(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    src = tmp_path / "c.cpp"
    src.write_text(PROG)
    exe = tmp_path / "c"
    libdir = os.path.dirname(LIB_PATH)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L", libdir, "-lesl", f"-Wl,-rpath,{libdir}"], check=True)
    return exe


def test_compat_header_host_parts(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe), model_path("toy.bnet")], capture_output=True, text=True, check=True).stdout
    assert out.split() == ["3", "2", "1", "0"]


@pytest.mark.gpu
def test_compat_header_solve(tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe), model_path("toy.bnet"), "gpu"], capture_output=True, text=True,
                         check=True).stdout.splitlines()
    t, a, b, c = out[1].split()
    # rates.toy golden (cuda/test/test.cu:556-558): 1/2, 11/48, 13/48
    assert t == "3" and abs(float(a) - 0.5) < 1e-5 and abs(float(b) - 11 / 48) < 1e-5 and abs(float(c) - 13 / 48) < 1e-5
