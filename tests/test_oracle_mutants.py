"""Planted mistakes in the oracle fail the reference's golden hashes, each in
the files of the stage it belongs to: builds of oracle/mp_oracle.c with
ORC_MUTANT set let the pair step join two reads only when both components
are small (&& for ||), drop a bit of the next character's class from the adj
byte, move the filter's ratio threshold from 0.2 to 0.25, or count 9-base
m-mer prefixes.  The golden comparison that pins the oracle is not blind to a
wrong logical operator, a wrong mask, a moved constant or a wrong length.
(Some mistakes leave the reference's output files unchanged -- which strand is
the canonical k-mer, prev and next swapped in the adj byte, a singleton tuple
kept at an 'N' -- and are pinned by the closed-form checks of
test_oracle_math.py instead: canonical k-mers, 'N' windows, neighbour bytes.)"""
import os
import shutil
import subprocess

import pytest

from _helpers import ROOT, diff_hashes, golden_case, run_chain

MUTANTS = [
    # ORC_MUTANT, golden case, files that must differ ("out": partitions only; "hist": histograms)
    (1, "c1_T2", "out"),
    (2, "c1_T2", "out"),
    (3, "c1_T2", "out"),
    (4, "c1_T2", "hist"),
]


@pytest.fixture(scope="module")
def mutant_builds(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")
    out = {}
    for n, _, _ in MUTANTS:
        exe = str(d / ("mp_oracle_m%d" % n))
        subprocess.run(["gcc", "-O2", "-std=c11", "-D_GNU_SOURCE", "-DORC_MUTANT=%d" % n, "-o", exe,
                        os.path.join(ROOT, "oracle", "mp_oracle.c"), os.path.join(ROOT, "oracle", "mp_oracle_main.c")],
                       check=True, stderr=subprocess.DEVNULL)
        out[n] = exe
    return out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("m", MUTANTS, ids=["mutant%d" % m[0] for m in MUTANTS])
def test_planted_mistake_fails_the_golden_hashes(m, mutant_builds, tmp_path):
    n, name, stage = m
    case = golden_case(name)
    try:
        got = run_chain("oracle", case, str(tmp_path), cli=mutant_builds[n])
    except RuntimeError:  # a mutant may also fail outright: caught all the same
        return
    bad = diff_hashes(got, case["sha256"])
    assert bad
    if stage == "out":  # index and histograms still match: the mistake is past them
        assert all(b.startswith("out/") for b in bad), bad
    else:
        assert any(b.startswith("KmerIdx") or b.startswith("kmerHist") for b in bad), bad
