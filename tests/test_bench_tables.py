"""CPU checks of bench.py's measurement plumbing: the committed HBM-traffic
table (profiles/evolve_traffic.json, from rocprofv3 FETCH_SIZE / WRITE_SIZE)
is self-consistent and feeds `roofline.traffic` for the launch lengths it
covers (fitted entries: any); the algorithmic flop count of a stepping launch matches the
operator's term count; `--gpus N` with N > 1 is refused."""
import json
import os
import subprocess
import sys

import pytest

from conftest import ROOT

sys.path.insert(0, ROOT)


def test_traffic_table_consistent():
    tab = json.load(open(os.path.join(ROOT, "profiles", "evolve_traffic.json")))["kernels"]
    assert tab
    for name, e in tab.items():
        fb, wb = e.get("fetch_bytes"), e.get("write_bytes")
        if e.get("resident"):
            # resident stepper: one figure per launch whatever its length (every
            # measured launch within 5 % of it), at least the state in and out once
            for la in e["launches"]:
                b = (la["fetch_KiB"] * 2048 + la["write_KiB"] * 1024) / e["parcels"]
                assert la["bytes_per_parcel"] == pytest.approx(b, abs=0.2)
                assert la["bytes_per_parcel"] == pytest.approx(e["bytes_per_parcel_launch"], rel=0.05)
            assert e["bytes_per_parcel_launch"] > 2 * 91 * 8
        elif "launches" in e:
            # fitted entry: every launch's bytes (copy-probe scale) within 2 % of
            # prologue + per-step x steps, and above the state in and out per step
            for la in e["launches"]:
                b = (la["fetch_KiB"] * 2048 + la["write_KiB"] * 1024) / e["parcels"]
                assert la["bytes_per_parcel"] == pytest.approx(b, abs=0.2)
                fit = e["prologue_bytes_per_parcel"] + e["bytes_per_parcel_step"] * la["steps"]
                assert fit == pytest.approx(b, rel=0.02)
            assert len({la["steps"] for la in e["launches"]}) >= 2
            assert e["bytes_per_parcel_step"] > 2 * 91 * 8
        elif name.startswith("evolve_fast"):
            # KiB counted -> bytes with the copy-probe scale (x 2048 read, x 1024 written)
            assert fb == pytest.approx(e["fetch_KiB"] * 2048, rel=1e-6)
            assert wb == pytest.approx(e["write_KiB"] * 1024, rel=1e-6)
            per = (fb + wb) / (e["parcels"] * e["steps_per_launch"])
            assert e["bytes_per_parcel_step"] == pytest.approx(per, abs=0.06)
            # at least the state in and out once per launch
            assert e["bytes_per_parcel_step"] * e["steps_per_launch"] > 2 * 91 * 8
        elif "bytes_per_parcel_launch" in e:
            # one launch per parcel batch (homogenization kernels): copy-probe scale
            assert fb == pytest.approx(e["fetch_KiB"] * 2048, rel=1e-6)
            assert wb == pytest.approx(e["write_KiB"] * 1024, rel=1e-6)
            assert e["bytes_per_parcel_launch"] == pytest.approx((fb + wb) / e["parcels"], abs=0.1)
        else:
            assert e["bytes_per_parcel"] > 0


def test_measured_traffic_lookup():
    """single-length entries serve their own launch length only; fitted
    entries any length (the driver's --steps 20 included)"""
    import bench
    tab = json.load(open(os.path.join(ROOT, "profiles", "evolve_traffic.json")))["kernels"]
    P = 1 << 20
    name, e = next((k, v) for k, v in tab.items()
                   if k.startswith("evolve_fast") and v.get("steps_per_launch") == 4)
    assert bench.measured_traffic(name, P, 4) == pytest.approx(e["bytes_per_parcel_step"] * P * 4)
    assert bench.measured_traffic(name, P, 3) is None          # launch length not measured
    f = tab["evolve_fast_L12_f64_v5"]
    for steps in (1, 10, 20, 200):
        assert bench.measured_traffic("evolve_fast_L12_f64_v5", P, steps) == pytest.approx(
            P * (f["prologue_bytes_per_parcel"] + f["bytes_per_parcel_step"] * steps))
    assert bench.measured_traffic("no_such_kernel", P, 10) is None
    r = next(v for k, v in tab.items() if v.get("resident"))
    rname = next(k for k, v in tab.items() if v is r)
    for steps in (1, 10, 200):
        assert bench.measured_traffic(rname, P, steps) == pytest.approx(P * r["bytes_per_parcel_launch"])


def test_flop_count_matches_operator():
    """72,214 flop per RK4 parcel-step at L = 12: 4 stages x (2 x 8,581
    coupling terms + 4 x 91 stage updates + 300 for <D>) + 10 x 91 for the
    RK combination; the term count is that of the generated stepper."""
    import bench
    assert bench.flops_per_parcel_step(12) == 4 * (2 * 8581 + 4 * 91 + 300) + 10 * 91 == 72214
    gen = os.path.join(ROOT, "specfab_amd", "csrc", "generated", "sfg_fast_L12_f64_v0.hip")
    if not os.path.exists(gen):
        pytest.skip("generated kernels not present (run the build)")
    # every coupling term is one FMA (or the chain's first multiply) on a literal
    lines = open(gen).read().splitlines()
    terms = sum(1 for line in lines if line.strip().startswith(("double u", "u")) and "x" in line)
    assert terms == 8581


def test_multi_gpu_request_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
