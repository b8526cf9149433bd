"""Detector error models: the circuit-to-DEM generator, the .dem parser and
the DEM TNMMAP network.

Pins: the reference's generator known answers (test/decoding/dem.jl:5-81),
reproducing the reference's dem.dem fixture from its noisy_circuit.stim
(test/decoding/dem.jl:83-109), the parse of dem.dem (test/decoding/
tndecoder.jl:112-120 uses it), and brute-force enumeration over all 2^21
mechanism subsets of dem.dem.
"""
import os

import numpy as np
import pytest

from oracle import tn_oracle as O
from oracle.cexec import OraclePlan
from tensorqec import decoders as D
from tensorqec.circuits import detector_error_model_from_stim, parse_stim, rotated_memory_z, to_dem_text
from tensorqec.dem import dem_error_patterns, parse_dem_file, parse_dem_string

HERE = os.path.join(os.path.dirname(__file__), "golden")
GOLDEN = np.load(os.path.join(HERE, "golden.npz"))


def as_set(dem):
    return {tuple(f): p for f, p in zip(dem.flipped_detectors, dem.error_rates)}


def test_generator_known_answer_depolarize_and_cx():
    dem = detector_error_model_from_stim("""X_ERROR(0.2) 0 1
DEPOLARIZE1(0.03) 0 1
CX 0 1
M 0 1
DETECTOR rec[-1]
DETECTOR rec[-2]
DETECTOR rec[-1] rec[-2]
OBSERVABLE_INCLUDE(0) rec[-1] rec[-2]
""")
    # dem.jl:22-23 (mechanism order there follows a Dict; compared as a set)
    got = as_set(dem)
    assert set(got) == {(1, 3, 4), (1, 2)}
    assert all(abs(p - 0.212) < 1e-10 for p in got.values())


def test_generator_known_answer_depolarize_only():
    dem = detector_error_model_from_stim("DEPOLARIZE1(0.3) 0\nM 0\nDETECTOR rec[-1]\n")
    assert dem.flipped_detectors == [[1]] and abs(dem.error_rates[0] - 0.2) < 1e-10   # dem.jl:36-37


def test_generator_known_answer_independent_components():
    dem = detector_error_model_from_stim("""H 0
CX 0 1
DEPOLARIZE1(0.0297) 0
H 3
CX 0 2 1 2 3 0 3 1
H 3
M 2 3
DETECTOR rec[-2]
DETECTOR rec[-1]
""")
    got = as_set(dem)                                                         # dem.jl:54-55
    assert set(got) == {(1,), (1, 2), (2,)}
    assert all(abs(p - 0.01) < 1e-10 for p in got.values())


def test_generator_known_answer_reset():
    dem = detector_error_model_from_stim("""X_ERROR(0.2) 0 1
CX 0 1
M 0 1
DETECTOR rec[-1] rec[-2]
R 1
X_ERROR(0.2) 0 1
CX 1 0
M 0 1
DETECTOR rec[-1] rec[-2]
""")
    got = as_set(dem)                                                         # dem.jl:79-80
    assert set(got) == {(1,), (2,)}
    assert abs(got[(1,)] - 0.2) < 1e-10 and abs(got[(2,)] - 0.32) < 1e-10


def test_generator_reproduces_reference_dem_fixture():
    # dem.jl:83-109: noisy_circuit.stim (stim color_code:memory_xyz) -> dem.dem
    d1 = detector_error_model_from_stim(open(os.path.join(HERE, "noisy_circuit.stim")).read())
    d2 = parse_dem_file(os.path.join(HERE, "dem.dem"))
    assert d1.detector_list == d2.detector_list and d1.logical_list == d2.logical_list
    a, b = as_set(d1), as_set(d2)
    assert set(a) == set(b)
    for k in a:
        assert abs(a[k] - b[k]) < 1e-10


def test_dem_text_round_trip():
    d1 = detector_error_model_from_stim(open(os.path.join(HERE, "noisy_circuit.stim")).read())
    d2 = parse_dem_string(to_dem_text(d1))
    assert as_set(d1).keys() == as_set(d2).keys()


def test_parser_rejects_repeat():
    with pytest.raises(ValueError):
        parse_dem_string("error(0.1) D0\nrepeat 2 {\n}\n")


def test_rotated_memory_circuit_structure():
    txt = rotated_memory_z(5, 5, 0.001)
    dem = detector_error_model_from_stim(txt)
    assert len(dem.detector_list) == 120 and dem.logical_list == [121]
    # every mechanism has a symptom; with no noise no mechanism exists
    assert all(len(f) >= 1 for f in dem.flipped_detectors)
    assert not np.any(detector_error_model_from_stim(rotated_memory_z(3, 2, 0.0)).error_rates)
    # a single round at d=3: 4 Z-check detectors + 4 final detectors
    assert len(detector_error_model_from_stim(rotated_memory_z(3, 1, 0.01)).detector_list) == 8


def test_rotated_memory_distance_3_needs_two_faults():
    """No single mechanism of the d=3 memory circuit flips the observable
    without flipping a detector (circuit distance > 1)."""
    dem = detector_error_model_from_stim(rotated_memory_z(3, 3, 0.001))
    L = dem.logical_list[0]
    assert all(f != [L] for f in dem.flipped_detectors)


@pytest.mark.parametrize("factorize,evidence", [(True, "slices"), (True, "leaves"), (False, "leaves")])
def test_dem_plan_matches_bruteforce(factorize, evidence):
    dem = parse_dem_file(os.path.join(HERE, "dem.dem"))
    ct = D.compile(D.TNMMAP(factorize=factorize, evidence=evidence), dem)
    m, cls, tie = OraclePlan(ct.plan).decode(GOLDEN["dem_syn"])
    ref = GOLDEN["dem_mar"]
    nz = ref > 0
    assert np.all(m[~nz] == 0)
    assert np.max(np.abs(m[nz] - ref[nz]) / ref[nz]) < 1e-12


def test_dem_error_pattern_rule():
    """tndecoder.jl:255-271: a pure error, then every mechanism of a logical
    whose parity disagrees with the class is flipped."""
    dem = parse_dem_file(os.path.join(HERE, "dem.dem"))
    ct = D.compile(D.TNMMAP(), dem)
    S = GOLDEN["dem_syn"]
    cls = np.arange(len(S)) % 2
    ep = dem_error_patterns(ct, S, cls)
    H = ct.tanner.H.astype(np.int64)
    pure = ((S.astype(np.int64) @ ct.pure.T.astype(np.int64)) & 1)
    assert np.array_equal((pure @ H.T) & 1, S)                 # the pure error is consistent
    mask = ct.l2q_mask[0].astype(np.int64)
    mismatch = ((pure @ mask) & 1) != cls
    expect = pure ^ (mismatch[:, None] * mask[None, :])
    assert np.array_equal(ep, expect)                           # the reference's literal rule
    # With an even number of mechanisms on the logical (6 here) the rule leaves
    # the parity unchanged; the reference's own test (tndecoder.jl:112-120)
    # only checks syndrome consistency, which the rule can break.
    assert mask.sum() == 6


def test_factorized_dem_refuses_single_member_checks():
    """tndecoder.jl:221-238 chains every check that does not have exactly two
    members through c[2], so a detector (or logical) flipped by a single
    mechanism fails there under `factorize`; the network builder refuses the
    same DEM with a ValueError naming the check, and builds it without
    factorization (one parity tensor per check), whose contraction still
    matches the brute-force marginal."""
    from tensorqec.network import tnmmap_dem_network
    text = ("error(0.1) D0 D1\nerror(0.2) D1 D2\nerror(0.05) D2 L0\nerror(0.3) D0 D2\n"
            "error(0.15) D3\nerror(0.25) D1 L0\n")
    dem = parse_dem_string(text)
    with pytest.raises(ValueError, match="detector 4 has 1 member"):
        tnmmap_dem_network(dem, factorize=True)
    g = tnmmap_dem_network(dem, factorize=False)
    assert len(g.ixs) == 4 + 1 + 6 + 4          # checks, logical, priors, evidence
    ct = D.compile(D.TNMMAP(factorize=False), dem)
    ns = 4
    rows = np.array([[(s >> j) & 1 for j in range(ns)] for s in range(1 << ns)], dtype=np.uint8)
    bf = O.brute_force_dem(*O.parse_dem_text(text))
    got, _, _ = OraclePlan(ct.plan).decode(rows)
    for r, row in zip(got, rows):
        want = bf.get(tuple(int(x) for x in row), np.zeros(2))
        assert np.allclose(r, want, rtol=1e-12, atol=1e-300)
    assert abs(got.sum() - 1.0) < 1e-12
