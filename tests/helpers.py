"""Shared test helpers: golden fixtures, model paths, oracle and product runs."""
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
MODELS = os.path.join(GOLDEN, "models")

# models of the reference's data/ directory kept as fixtures
REAL_MODELS = ["toy.bnet", "toy2.bnet", "toy3.bnet", "toy-paper.bnet", "krasmodel15vars.bnet",
               "mammalian_cc.bnet", "EMT_cohen_ModNet.bnet", "zanudo_expanded.bnet"]
SMALL_DB = sorted(os.listdir(os.path.join(MODELS, "database-small")))


def model_path(name):
    return os.path.join(MODELS, name)


def golden():
    with open(os.path.join(GOLDEN, "reference_tests.json")) as f:
        g = json.load(f)
    return {k: v for k, v in g.items() if not k.startswith("_")}


def oracle_rates(model, case, key="rates"):
    from oracle import pipeline as P

    r = case.get(key, {})
    return P.rates_uniform(model, r.get("up"), r.get("down"))


def product_rates(model, case, key="rates"):
    from paraexastolog_amd import transition_rates

    r = case.get(key, {})
    tr = transition_rates(model)
    tr.generate_uniform(list(r.get("up", {}).items()), list(r.get("down", {}).items()))
    return tr


def nonzero(x):
    nz = np.nonzero(x)[0]
    return nz, x[nz]


def l1(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum())
