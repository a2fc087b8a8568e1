"""The oracle against its committed self-golden answers (CPU).

tests/golden/self_golden_answers.json holds answers the oracle produced once
(tests/golden/make_answer_digests.py) for the reference's own e2e seeds and
for the toy n1 = 64 VIA-C stack at 2^10 records; the GPU tests compare the
HIP path's bytes against the same file. Recomputing them here keeps the file
and the oracle in step: a change to the oracle that moves any answer word
fails here, not only on the GPU.
"""
import json
import os

import numpy as np
import pytest

from tests.golden import make_answer_digests as mk

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "self_golden_answers.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def test_toy_n8_every_index(oracle, golden):
    # client_server_e2e.rs:92-172 (each answer also recovers its record)
    assert mk.toy_n8_all(oracle)["answers"] == golden["toy_n8_all"]["answers"]


def test_toy_b_n64_t8(oracle, golden):
    # batch_e2e_toy.rs:190-225, first index set
    got = mk.toy_b_n64(oracle)
    want = golden["toy_b_n64_t8"]
    assert (got["mask"], got["body"], got["indices"]) == (want["mask"], want["body"], want["indices"])


def test_toy_n64_via_c_1k_records(oracle, golden):
    # the n1 = 64 single-prime stack of batch_e2e_toy.rs:40-68 at the VIA-C
    # record degree: 2^10 seeded records, I = J = 16, 70 query LWEs
    got = mk.toy_n64_1k(oracle)
    want = golden["toy_n64_1k_records"]
    assert got["query_lwes"] == 70 == want["query_lwes"]
    g, w = got["answers"][0], want["answers"][0]
    assert g == w
    assert g["recovered"] == g["record"]


@pytest.mark.slow
def test_paper_2x2_index15(oracle, golden):
    # client_server_e2e_paper.rs:287-349
    assert mk.paper_2x2_15(oracle)["answers"] == golden["paper_2x2_index15"]["answers"]


def test_fixture_cases_present(golden):
    # the full-size digests the GPU tests compare against
    for name in ("paper_128x128", "paper_256x256", "paper_512x512", "vb_t8_32x4096", "vb_t16_32x4096",
                 "vb_t32_32x4096", "vb_t64_32x4096"):
        c = golden[name]
        assert c["sha256"] and all(len(h) == 64 for h in c["sha256"]), name
        assert c["indices"][0] == 0
    assert np.unique([len(a["mask"]) for a in golden["toy_n8_all"]["answers"]]).tolist() == [4]
