"""The Node recorder restatement (oracle/recorder.py) against the reference's recorder output.

- The due rule and the time accumulation: the time column of the reference's
  column25_ssp_explicit case (OpenSees, `-dT 0.01` at dt 2.5e-4, precision 17) is reproduced
  token for token by accumulating the step and applying the rule.
- The line format: the first lines of column4_pimy_std_implicit/opensees_disp.txt, parsed and
  printed again, give the same text (%.17g of each value, fixed DOFs printed 0).
- Plausible mistakes in the rule (no tolerance, next advanced from the last due time instead of
  the record time, the step counted from the start) or in the format (%.16g, %e) give other
  text, so the fixtures tell them apart.
"""
import json
import os

import pytest

from oracle import recorder as R

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "recorder_text.json")


@pytest.fixture(scope="module")
def text():
    with open(FIXTURE) as f:
        return json.load(f)


def times_text(times, idx, precision=17):
    return [f"{times[i]:.{precision}g}" for i in idx]


def test_due_rule_reproduces_the_recorded_times(text):
    c = text["column25_ssp_explicit_times"]
    times = R.accumulated_times(c["dt"], 8000)
    got = times_text(times, R.due_indices(times, c["dT"]))[:len(c["times"])]
    assert got == c["times"]


def test_planted_rule_mistakes_change_the_times(text):
    c = text["column25_ssp_explicit_times"]
    times = R.accumulated_times(c["dt"], 8000)
    want = c["times"]

    def strict(ts, dT):  # no tolerance
        out, nxt = [], 0.0
        for i, t in enumerate(ts):
            if t - nxt >= 0.0:
                nxt = t + dT
                out.append(i)
        return out

    def from_next(ts, dT):  # next advanced from itself
        out, nxt = [], 0.0
        for i, t in enumerate(ts):
            if t - nxt >= -dT * 1e-5:
                nxt = nxt + dT
                out.append(i)
        return out

    for rule in (strict, from_next):
        assert times_text(times, rule(times, c["dT"]))[:len(want)] != want
    # times computed as k dt instead of accumulated
    product = [(k + 1) * c["dt"] for k in range(8000)]
    assert times_text(product, R.due_indices(product, c["dT"]))[:len(want)] != want


def test_line_format_reproduces_opensees_lines(text):
    c = text["column4_pimy_std_implicit_disp_head"]
    for line in c["lines"]:
        vals = [float(x) for x in line.split()]
        assert R.line(vals[0], vals[1:], precision=c["precision"]) == line
        assert R.line(vals[0], vals[1:], precision=16) != line
        assert " ".join(f"{v:.16e}" for v in vals) + "\n" != line
    assert R.line(0.5, [0.0, -0.0, 1.5], precision=6, echo_time=False) == "0 -0 1.5\n"


def test_every_step_and_the_first_commit():
    times = R.accumulated_times(0.005, 10)
    assert R.due_indices(times, 0.0) == list(range(10))
    # dT = 2 dt: the first commit, then every second one
    assert R.due_indices(times, 0.01) == [0, 2, 4, 6, 8]
    # dT below dt records every commit
    assert R.due_indices(times, 0.004) == list(range(10))
