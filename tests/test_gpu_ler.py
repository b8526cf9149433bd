"""Logical-error-rate loop on the device (tqec_ler_*) against the host
restatement of the reference harness, through the C ABI.

Reference loop: `random_error_pattern` -> `syndrome_extraction` -> `decode`
-> `check_logical_error` (src/decoding/error_model.jl:97-181; the loops of
test/decoding/tndecoder.jl:85-99 and src/decoding/threshold.jl:1-32).  The
host side samples the same Philox stream (bit-identical), decodes with the
same plan and mode, recovers error patterns with the host restatement of
`error_pattern` and applies `check_logical_error`; the four counts (logical
errors, X half, Z half, recovered patterns that miss the syndrome) must be
equal exactly.  Splitting the samples over two calls gives the same totals.
"""
import os
import sys

import numpy as np
import pytest

from tensorqec import decoders as D
from tensorqec import lib as L
from tensorqec.codes import CSSTannerGraph, SurfaceCode, ToricCode
from tensorqec.dem import parse_dem_file, random_error_pattern
from tensorqec.errors import (check_logical_error, check_logical_error_H, css_syndrome_rows, iid_error,
                              sample_depolarizing, syndrome_extraction)

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("code,p,mode", [(SurfaceCode(5, 5), 0.05, "spec_f32"), (SurfaceCode(7, 7), 0.08, "spec_f32"),
                                         (ToricCode(3, 3), 0.06, "plain_f64")])
def test_css_ler_loop_matches_host(code, p, mode):
    t = CSSTannerGraph.from_code(code)
    em = iid_error(p / 3, p / 3, p / 3, t.nq)
    ct = D.compile(D.TNMMAP(mode=mode), t, em)
    smp = L.HipSampler.css(t, em, ct.lx, ct.lz)
    rec = L.HipRecovery.css(t, ct.pure_x, ct.pure_z, ct.lx, ct.lz)
    batch = 4096
    ler = L.HipLER(ct.hip, smp, rec, batch=batch)
    seed, first, n = 20261203, 777, 3 * batch + 1234
    say(f"case ler {code} p={p} {mode}: {n} samples")
    got = ler.run(mode, seed, first, n)
    ep = sample_depolarizing(em, seed, n, first=first)
    S = css_syndrome_rows(ep, t)
    r = D.decode_batch(ct, S, mode=mode)
    lerr = check_logical_error(ep, r.error_patterns, ct.lx, ct.lz)
    xh = check_logical_error_H(ep.xerror, r.error_patterns.xerror, ct.lz)
    zh = check_logical_error_H(ep.zerror, r.error_patterns.zerror, ct.lx)
    sx, sz = syndrome_extraction(r.error_patterns, t)
    bad = (np.concatenate([sx, sz], axis=1) != S).any(axis=1)
    want = {"samples": n, "logical_errors": int(lerr.sum()), "x_half": int(xh.sum()), "z_half": int(zh.sum()),
            "inconsistent": int(bad.sum())}
    say(f"  device {got} host {want}")
    assert got == want
    assert want["inconsistent"] == 0 and want["logical_errors"] > 0
    # the same samples over two calls
    a = ler.run(mode, seed, first, 5000)
    b = ler.run(mode, seed, first + 5000, n - 5000)
    for k in ("logical_errors", "x_half", "z_half", "inconsistent"):
        assert a[k] + b[k] == got[k]


def test_dem_ler_loop_matches_host():
    dem = parse_dem_file(os.path.join(HERE, "dem.dem"))
    ct = D.compile(D.TNMMAP(mode="fused_f32"), dem)
    smp = L.HipSampler.flips(dem.error_rates, dem.detector_matrix(), dem.logical_matrix())
    rec = L.HipRecovery.dem(ct.pure, ct.l2q_mask)
    ler = L.HipLER(ct.hip, smp, rec, batch=1000)
    seed, n = 99, 4321
    got = ler.run("fused_f32", seed, 0, n)
    e = random_error_pattern(dem, seed, n)
    H = ct.tanner.H.astype(np.int64)
    S = ((e.astype(np.int64) @ H.T) & 1).astype(np.uint8)
    r = D.decode_batch(ct, S)
    lerr = check_logical_error_H(e, r.error_patterns, dem.logical_matrix())
    bad = (((r.error_patterns.astype(np.int64) @ H.T) & 1) != S).any(axis=1)
    want = {"samples": n, "logical_errors": int(lerr.sum()), "x_half": 0, "z_half": 0, "inconsistent": int(bad.sum())}
    say(f"case ler dem.dem: device {got} host {want}")
    assert got == want


def test_ler_rejects_mismatched_parts():
    t5 = CSSTannerGraph.from_code(SurfaceCode(5, 5))
    t3 = CSSTannerGraph.from_code(SurfaceCode(3, 3))
    em5 = iid_error(0.01, 0.01, 0.01, t5.nq)
    em3 = iid_error(0.01, 0.01, 0.01, t3.nq)
    ct5 = D.compile(D.TNMMAP(), t5, em5)
    ct3 = D.compile(D.TNMMAP(), t3, em3)
    smp3 = L.HipSampler.css(t3, em3, ct3.lx, ct3.lz)
    rec5 = L.HipRecovery.css(t5, ct5.pure_x, ct5.pure_z, ct5.lx, ct5.lz)
    with pytest.raises(L.TqecError) as e:
        L.HipLER(ct5.hip, smp3, rec5)
    assert e.value.code == L.TQEC_EINVAL
