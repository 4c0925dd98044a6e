"""The two KERN-05 seahorse samples that the frozen bundle classes Interior and the code as
written (sample.rs PERF-08), the oracle and the GPU all class Unresolved: the trace.

Samples 765 and 12093 of the seahorse frame (crates/fd-kernel/fixtures/KERN-05.bundle,
-0.7453 + 0.1127i, width 0.0065, 192x108, max_iter 20,000, f64 tier) settle on attracting
cycles of 162 = 2 x 81 and 232 = 4 x 58 steps.  The Brent/Newton confirmation
(sample.rs:524-558) sees a near-return every 81 (58) steps, measures the return
multiplier mu over each, and predicts the cycle point only when |mu| < 1; Interior needs
two consecutive predictions that agree.  From one near-return to the next |mu| alternates
above and below 1 (the whole cycle contracts: the product over it is below 1), so every
prediction is followed by a return that makes none: the sample runs to max_iter.
oracle/trace.py restates the loop with a log of every near-return; it is first checked
against the oracle's own outcomes.
"""
import numpy as np
import pytest

import oracle
from fdsfile import kern05, parse_fds
from oracle.trace import sample_plain_f64

# sample index: (steps between near-returns, near-returns per cycle)
CASES = {765: (81, 2), 12093: (58, 4)}


@pytest.fixture(scope="module")
def frame():
    h, want = parse_fds(dict(kern05())["seahorse"])
    assert h["kernel"] == "pert-f64/1"
    pl = oracle.plane(h["re"], h["im"], h["width"], h["rotation"], h["nx"], h["ny"])
    zr, zi, _ = oracle.reference(h["re"], h["im"], h["width"], h["rotation"], h["nx"],
                                 h["ny"], h["max_iter"], h["escape_radius"])
    nuc = oracle.nucleus_period(zr, zi)
    hh = pl["h_m"] * 2.0 ** pl["h_e"]

    def offset(idx):
        # view.rs:117-122 unit_offset, times h
        x = idx % h["nx"] + 0.5 - pl["half_x"]
        y = idx // h["nx"] + 0.5 - pl["half_y"]
        return ((pl["cos_r"] * x + pl["sin_r"] * y) * hh,
                (pl["sin_r"] * x - pl["cos_r"] * y) * hh)

    return dict(h=h, want=want, pl=pl, zr=zr, zi=zi, nuc=nuc, offset=offset,
                r2=h["escape_radius"] ** 2)


def run(f, idx, log=None):
    ar, ai = f["offset"](idx)
    c = (f["pl"]["c_re"] + ar, f["pl"]["c_im"] + ai)
    return sample_plain_f64(f["zr"], f["zi"], c, ar, ai, f["h"]["max_iter"], f["r2"], f["nuc"],
                            log)


def test_trace_restatement_equals_the_oracle(frame):
    # every sample of the two rows around sample 765, and both traced samples
    kinds = {0: "escaped", 1: "interior", 2: "unresolved"}
    idxs = list(range(576, 960)) + list(CASES)
    got_kinds = set()
    for idx in idxs:
        ar, ai = frame["offset"](idx)
        c = (frame["pl"]["c_re"] + ar, frame["pl"]["c_im"] + ai)
        o = oracle.sample(frame["zr"], frame["zi"], ar, ai, frame["h"]["max_iter"], frame["r2"],
                          c=c, nucleus=frame["nuc"])
        want = (kinds[o["kind"]], o["n"] if o["kind"] != 2 else 0)
        got = run(frame, idx)
        assert got == want, (idx, got, want)
        got_kinds.add(got[0])
    assert got_kinds == {"escaped", "interior", "unresolved"}


@pytest.mark.parametrize("idx", sorted(CASES))
def test_return_multipliers_alternate_so_no_prediction_is_confirmed(frame, idx):
    half, k = CASES[idx]
    # the bundle says Interior; the code as written says Unresolved
    assert frame["want"]["class"][idx] & 3 == 1
    log = []
    assert run(frame, idx, log) == ("unresolved", 0)
    # past the first few returns the orbit has settled: a near-return every `half` steps
    settled = [e for e in log if e["n"] > 2000]
    assert len(settled) > 100
    assert all(e["period"] == half and e["repeat"] for e in settled)
    below = [e["below"] for e in settled]
    # |mu| from one near-return to the next alternates above and below 1 ...
    assert all(a != b for a, b in zip(below, below[1:]))
    # ... so a prediction (made at a return with |mu| < 1) never meets a second one: the
    # return after it has |mu| > 1 and drops it
    for a, b in zip(settled, settled[1:]):
        if a["below"]:
            assert not b["below"] and b["predicted"]
        else:
            assert not b["predicted"]
    # the multipliers repeat with the cycle (k returns), whose product contracts: |mu| < 1
    mus = np.array([e["log2_mu"] for e in settled])
    assert np.all(np.abs(mus[k:] - mus[:-k]) < 0.02)
    assert all(mus[i:i + k].sum() < -0.5 for i in range(len(mus) - k))
    # and the orbit repeats with the cycle: returns k apart agree far more closely than
    # consecutive ones
    z = np.array([e["z"] for e in settled])
    one = np.abs(z[1:] - z[:-1]).sum(axis=1)
    cyc = np.abs(z[k:] - z[:-k]).sum(axis=1)
    assert np.median(cyc) < 0.2 * np.median(one)
