"""Host-side compile logic: codes, GF(2) algebra, logical operators, sampler.

Pins: test/codes/code_distance.jl:54-59 (d=3 lx/lz golden), test/codes/codes.jl:5-40
(toric stabilizers and bimatrix goldens), test/decoding/tndecoder.jl:8-14
(parity tensor entries), Random123's Philox4x32-10 known-answer vectors.
"""
import os

import numpy as np
import pytest

from tensorqec.codes import (CSSTannerGraph, ShorCode, SteaneCode, SurfaceCode, ToricCode, stabilizers,
                             stabilizers2bimatrix)
from tensorqec.errors import (check_logical_error, css_syndrome_rows, iid_error, philox4x32, sample_depolarizing,
                              sample_flips, syndrome_extraction_H, CSSErrorPattern)
from tensorqec.gf2 import gf2_rank, gf2_right_inverse, logical_operator, null_space
from tensorqec.network import parity_check_tensor, qubit_prior

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))

CONFIG_CODES = {f"surface{d}": SurfaceCode(d, d) for d in (3, 5, 7, 9)}
CONFIG_CODES.update({f"toric{L}": ToricCode(L, L) for L in (2, 3, 4, 5, 6)})
CONFIG_CODES.update({"steane": SteaneCode(), "shor": ShorCode()})


def test_surface3_logical_golden():
    lx, lz = logical_operator(CSSTannerGraph.from_code(SurfaceCode(3, 3)))
    assert lx.tolist() == [[0, 0, 0, 0, 0, 0, 1, 1, 1]]
    assert lz.tolist() == [[1, 0, 0, 0, 1, 0, 0, 0, 1]]


def test_surface3_stabilizer_sets():
    # src/codes/codes.jl:24-26: X type 36, 1245, 5689, 47; Z type 12, 2356, 4578, 89
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    xs = sorted(tuple(q + 1 for q in c) for c in t.stgx.s2q)
    zs = sorted(tuple(q + 1 for q in c) for c in t.stgz.s2q)
    assert xs == sorted([(3, 6), (1, 2, 4, 5), (5, 6, 8, 9), (4, 7)])
    assert zs == sorted([(1, 2), (2, 3, 5, 6), (4, 5, 7, 8), (8, 9)])


def test_toric23_golden():
    exp = [(2, 1, 2, 1, 1, 1, 2, 2, 1, 1, 1, 1), (1, 2, 1, 2, 1, 1, 2, 2, 1, 1, 1, 1),
           (1, 1, 2, 1, 2, 1, 1, 1, 2, 2, 1, 1), (1, 1, 1, 2, 1, 2, 1, 1, 2, 2, 1, 1),
           (2, 1, 1, 1, 2, 1, 1, 1, 1, 1, 2, 2), (4, 4, 1, 1, 1, 1, 4, 1, 1, 1, 4, 1),
           (4, 4, 1, 1, 1, 1, 1, 4, 1, 1, 1, 4), (1, 1, 4, 4, 1, 1, 4, 1, 4, 1, 1, 1),
           (1, 1, 4, 4, 1, 1, 1, 4, 1, 4, 1, 1), (1, 1, 1, 1, 4, 4, 1, 1, 4, 1, 4, 1)]
    st = stabilizers(ToricCode(2, 3))
    assert len(st) == len(exp)
    for a, e in zip(st, exp):
        assert a.tolist() == [x - 1 for x in e]


def test_toric22_bimatrix_golden():
    b = stabilizers2bimatrix(stabilizers(ToricCode(2, 2)))
    assert b.xcodenum == 3
    assert b.ordering == list(range(8))
    assert b.matrix.tolist() == [
        [1, 0, 1, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
        [0, 1, 0, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0],
        [1, 0, 1, 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0],
        [0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 1, 0, 1, 0],
        [0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 0, 0, 0, 1, 0, 1],
        [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 0, 1, 0]]


@pytest.mark.parametrize("name", sorted(CONFIG_CODES))
def test_logical_invariants(name):
    t = CSSTannerGraph.from_code(CONFIG_CODES[name])
    lx, lz = logical_operator(t)
    Hx, Hz = t.stgx.H.astype(int), t.stgz.H.astype(int)
    k = t.nq - gf2_rank(Hx) - gf2_rank(Hz)
    assert lx.shape[0] == lz.shape[0] == k
    assert not ((Hx @ lz.T) % 2).any()          # Z logicals commute with X checks
    assert not ((Hz @ lx.T) % 2).any()
    assert np.array_equal((lx.astype(int) @ lz.T.astype(int)) % 2, np.eye(k, dtype=int))
    assert gf2_rank(np.vstack([Hx, lx])) == gf2_rank(Hx) + k   # not stabilizers
    assert gf2_rank(np.vstack([Hz, lz])) == gf2_rank(Hz) + k
    # regression against the committed rows
    assert np.array_equal(lx, GOLDEN[f"lx_{name}"]) and np.array_equal(lz, GOLDEN[f"lz_{name}"])


def test_null_space():
    rng = np.random.default_rng(3)
    for _ in range(20):
        H = rng.integers(0, 2, size=(6, 11)).astype(np.uint8)
        K = null_space(H)
        assert K.shape[0] == 11 - gf2_rank(H)
        assert not ((H.astype(int) @ K.T.astype(int)) % 2).any()
        assert gf2_rank(K) == K.shape[0]


def test_parity_check_tensor_entries():
    # test/decoding/tndecoder.jl:8-14 (Julia indices are 1-based, first index fastest)
    t = parity_check_tensor(4)

    def at(*a):
        return t[sum((x - 1) << j for j, x in enumerate(a))]

    assert at(1, 1, 1, 1, 1) == 1 and at(1, 1, 1, 1, 2) == 0
    assert at(1, 2, 1, 2, 1) == 1 and at(2, 1, 2, 2, 1) == 0


def test_qubit_prior_layout():
    # [1-px-py-pz pz; px py] indexed [x, z], column-major (tndecoder.jl:118-121)
    t = qubit_prior(0.1, 0.2, 0.3)
    assert t.tolist() == pytest.approx([0.4, 0.1, 0.3, 0.2])


def test_pure_error_map():
    rng = np.random.default_rng(5)
    for code in (SurfaceCode(5, 5), ToricCode(4, 4), SteaneCode(), ShorCode()):
        t = CSSTannerGraph.from_code(code)
        for H in (t.stgx.H, t.stgz.H):
            R = gf2_right_inverse(H)
            e = rng.integers(0, 2, size=(50, t.nq)).astype(np.uint8)
            s = syndrome_extraction_H(e, H)
            e2 = (s.astype(int) @ R.T.astype(int)) % 2
            assert np.array_equal(syndrome_extraction_H(e2.astype(np.uint8), H), s)


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, e in kat:
        assert tuple(int(x) for x in philox4x32(*c, *k)) == e


def test_sampler_statistics_and_semantics():
    n = 49
    em = iid_error(0.02, 0.03, 0.05, n)
    ep = sample_depolarizing(em, seed=11, count=20000)
    y = ep.xerror & ep.zerror
    x = ep.xerror & (1 - ep.zerror)
    z = ep.zerror & (1 - ep.xerror)
    for arr, p in ((x, 0.02), (y, 0.03), (z, 0.05)):
        f = arr.mean()
        assert abs(f - p) < 5 * np.sqrt(p * (1 - p) / arr.size)
    # determinism and counter addressing: a slice equals the same rows of a larger draw
    ep2 = sample_depolarizing(em, seed=11, count=100, first=500)
    assert np.array_equal(ep2.xerror, ep.xerror[500:600]) and np.array_equal(ep2.zerror, ep.zerror[500:600])
    f = sample_flips(iid_error(0.1, 30), seed=2, count=20000)
    assert abs(f.mean() - 0.1) < 0.005


def test_syndrome_and_logical_check():
    t = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    lx, lz = logical_operator(t)
    ep = CSSErrorPattern(np.array([1, 0, 0, 0, 0, 0, 0, 0, 0], np.uint8), np.zeros(9, np.uint8))
    rows = css_syndrome_rows(ep, t)
    assert rows.shape == (1, 8)
    # X1 flips the Z check on qubits (1,2) only
    assert rows[0, :4].sum() == 0 and rows[0, 4:].tolist() == [0, 0, 1, 0]
    ep2 = CSSErrorPattern(ep.xerror ^ lx[0], ep.zerror)
    assert bool(check_logical_error(ep, ep2, lx, lz))
