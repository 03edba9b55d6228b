"""The FP32 -> FP64 hand-off policy of the MIXED mode (csrc/handoff.hpp), on
the CPU: a small driver compiled with the host C++ compiler feeds it residue
sequences with known rates and floors and reads its decisions back."""
import math
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "gaugefix-coulomb_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("handoff") / "handoff_driver")
    subprocess.run([cxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                    os.path.join(ROOT, "tests", "handoff_driver.cpp"), "-o", exe], check=True)
    return exe


def run(driver, lines):
    out = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True,
                         text=True, check=True).stdout.split("\n")
    res = []
    for l in out:
        if l.strip():
            r, f, fl, sw, rho = l.split()
            res.append(dict(ret=int(r), finished=int(f), floor=int(fl), switch_at=int(sw),
                            rho=float(rho)))
    return res


def geometric(r0, decades_per_sweep, every, n):
    """Probe residues of a geometric convergence: r0 * 10^(-rate * sweep)."""
    return [(k * every, r0 * 10 ** (-decades_per_sweep * k * every)) for k in range(1, n + 1)]


def test_rate_and_switch_on_a_geometric_sequence(driver):
    rate, due, tol = 0.02, 1085, 1e-16
    seq = geometric(16.0, rate, 50, 12)  # down to 16e-12 at sweep 600
    lines = ["O 1e-9 1 1e-16", "S 16"] + [f"X {s} {g!r} {due}" for s, g in seq]
    res = run(driver, lines)
    for (s, g), r in zip(seq, res):
        assert r["rho"] == pytest.approx(rate, rel=1e-12)
        assert r["floor"] == 0
        assert r["finished"] == (g <= 1e-9)
        if not r["finished"]:
            assert r["ret"] == 0 and r["switch_at"] == 0x7fffffff
            continue
        # deferred: the FP64 sweeps left before the check reach the tolerance at
        # the measured rate, with a bounded margin
        left = due - r["switch_at"]
        need = math.log10(g / tol) / rate
        assert r["ret"] == 0 and s < r["switch_at"] < due
        assert need <= left <= 1.1 * need + 11


def test_immediate_handoff_without_deferral(driver):
    seq = geometric(16.0, 0.02, 50, 12)
    res = run(driver, ["O 1e-9 0 1e-16", "S 16"] + [f"X {s} {g!r} 1085" for s, g in seq])
    first = next(i for i, (s, g) in enumerate(seq) if g <= 1e-9)
    assert all(r["ret"] == 0 for r in res[:first])
    assert res[first]["ret"] == 1


def test_floor_is_detected_and_stops_probing(driver):
    # converging at 0.02 decades/sweep to 1.6e-9 at sweep 500, slowing down
    # (0.5 of the rate), then flat at the FP32 rounding floor
    seq = geometric(16.0, 0.02, 50, 10) + [(550, 5e-10), (600, 4.9e-10), (650, 4.95e-10)]
    res = run(driver, ["O 0 1 1e-16", "S 16"] + [f"X {s} {g!r} 100000" for s, g in seq])
    assert all(r["floor"] == 0 for r in res[:10])
    # the slow-down neither counts as a clean interval nor as the floor
    assert res[10]["floor"] == 0 and res[10]["finished"] == 0
    assert res[10]["rho"] == pytest.approx(0.02, rel=1e-12)
    assert res[11]["floor"] == 1 and res[11]["finished"] == 1
    assert res[11]["rho"] == pytest.approx(0.02, rel=1e-12)
    assert res[11]["ret"] == 0 and 600 < res[11]["switch_at"] < 100000
    # at the floor the rate is not updated any more
    assert res[12]["floor"] == 1 and res[12]["rho"] == res[11]["rho"]


def test_handoff_when_the_check_is_too_close(driver):
    seq = geometric(16.0, 0.02, 50, 11)  # reaches 1.6e-10 at sweep 550
    s, g = seq[-1]
    need = math.log10(g / 1e-16) / 0.02
    due = s + int(need)  # fewer sweeps left than the FP64 phase needs
    res = run(driver, ["O 1e-9 1 1e-16", "S 16"] + [f"X {a} {b!r} {due}" for a, b in seq])
    assert res[-1]["finished"] == 1 and res[-1]["ret"] == 1


def test_below_tolerance(driver):
    # a loose tolerance above mixed_switch: once the FP32 residue is at or below
    # both, the hand-off is immediate (no FP64 sweeps are needed)
    res = run(driver, ["O 1e-5 1 1e-6", "S 1", "X 50 1e-3 1000", "X 100 5e-7 1000"])
    assert res[0]["ret"] == 0
    assert res[1]["ret"] == 1
    # below the tolerance but above mixed_switch the FP32 phase goes on: the
    # next scheduled check can pass in FP32
    res = run(driver, ["O 1e-9 1 1e-6", "S 1", "X 50 1e-3 1000", "X 100 5e-7 1000"])
    assert res[1]["ret"] == 0 and res[1]["finished"] == 0
