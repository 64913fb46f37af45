"""Pins the renderer pieces of the oracle: Philox against the Random123
known answers, the sin/cos polynomial against libm, the alias table, the
roulette estimator and bajo's wavefront contract constants; checks that the
library's host view of the stage streams equals the oracle's.  CPU only."""

import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_vectors.json").read_text())


def test_philox_known_answers():
    for ctr, key, want in GOLD["philox4x32_10_kat"]["cases"]:
        got = oracle.philox4x32_10([int(x, 16) for x in ctr], [int(x, 16) for x in key])
        assert [f"{int(x):08x}" for x in got] == want


def test_sincos_turns_against_libm():
    u = np.concatenate([np.linspace(0, 1, 20001, endpoint=False, dtype=np.float32), np.float32([0.125, 0.375, 0.625, 0.875])])
    worst = 0.0
    for x in u[::7]:
        s, c = oracle.sincos_turns(float(x))
        ang = 2.0 * math.pi * float(x)
        worst = max(worst, abs(s - math.sin(ang)), abs(c - math.cos(ang)))
        assert s * s + c * c == pytest.approx(1.0, abs=3e-7)
    assert worst < 2.5e-7  # within ~2 ulp of 1.0


def test_stage_streams_are_uniform_and_separate():
    f = oracle.rng_stage_floats(91, 7, 3, 4096)
    assert np.all((f >= 0) & (f < 1))
    assert abs(f.mean() - 0.5) < 0.02
    replay = oracle.rng_stage_floats(91, 7, 3, 8)
    np.testing.assert_array_equal(f[:8], replay)
    assert not np.array_equal(oracle.rng_stage_floats(91, 7, 4, 8), replay)
    assert not np.array_equal(oracle.rng_stage_floats(91, 7, (1 << 30) | 3, 8), replay)
    assert not np.array_equal(oracle.rng_stage_floats(92, 7, 3, 8), replay)


def test_library_host_stream_equals_oracle():
    from bajo_amd import rt

    for seed, pid, stage in ((0, 0, 0), (2026, 123456, rt.wavefront_rng_light_stage(3)), (2**40 + 5, 2**31 - 1, 7)):
        np.testing.assert_array_equal(rt.rng_stage_floats(seed, pid, stage, 11), oracle.rng_stage_floats(seed, pid, stage, 11))


def test_wavefront_contract_constants():
    from bajo_amd import rt

    g = GOLD["wavefront_contract"]
    assert rt.wavefront_rng_subsequence(7, 3) == g["subsequence_7_3"]
    assert rt.wavefront_rng_stage(0) == g["bsdf_stage_0"]
    assert rt.wavefront_rng_stage(6) == g["bsdf_stage_6"]
    assert rt.wavefront_rng_roulette_stage(0) == g["roulette_stage_0"]
    assert rt.wavefront_rng_light_stage(0) == g["light_stage_0"]
    assert rt.wavefront_rng_roulette_stage(6) != rt.wavefront_rng_stage(6)
    assert rt.wavefront_rng_light_stage(6) != rt.wavefront_rng_roulette_stage(6)


def test_alias_table_matches_power_distribution():
    g = GOLD["alias_table"]
    prob, alias = oracle.alias_table(g["weights"])
    rec = np.zeros(3)
    for col in range(3):
        assert 0 <= prob[col] <= 1 and 0 <= alias[col] < 3
        rec[col] += prob[col] / 3
        rec[alias[col]] += (1 - prob[col]) / 3
    np.testing.assert_allclose(rec, g["expected"], atol=g["atol"])


def test_alias_table_random_weights_reconstruct():
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 64):
        w = rng.uniform(0.01, 5.0, n).astype(np.float32)
        prob, alias = oracle.alias_table(w)
        rec = np.zeros(n)
        for col in range(n):
            rec[col] += prob[col] / n
            rec[alias[col]] += (1 - prob[col]) / n
        np.testing.assert_allclose(rec, w / w.sum(), atol=2e-6)


def test_light_table_weights_are_area_times_mean_radiance():
    from bajo_amd import scenes

    sc = scenes.box_scene()
    fields, kinds, total = oracle.light_table(sc.vertices, sc.surfaces, sc.emissive)
    assert len(kinds) == 2 and np.all(kinds & 0xF == 1)
    # two triangles of a 0.6 x 0.6 quad, radiance 12
    assert total == pytest.approx(0.36 * 12.0, rel=1e-6)


def test_roulette_is_unbiased():
    # rt/common.mojo:43-67 through the oracle's roulette stage stream
    g = GOLD["roulette"]
    t = np.float32(g["throughput"][0])
    survival = min(max(max(g["throughput"]), 0.05), 0.95)
    acc = 0.0
    for pid in range(g["trials"]):
        u = oracle.rng_stage_floats(8128, pid, (1 << 30) | 5, 1)[0]
        if u < survival:
            acc += float(np.float32(t / np.float32(survival)))
    assert abs(acc / g["trials"] - t) < g["atol"]


def test_fp32_oracle_agrees_with_fp64_estimator():
    """The two independent estimators of the same image agree statistically."""
    from bajo_amd import scenes
    from bajo_amd.bvh import Camera

    sc = scenes.box_scene()
    cam = Camera((0.0, 1.0, -3.0), (0.0, 0.9, 0.0), (0.0, 1.0, 0.0), 0.5).flatten()
    for alg in (0, 3):
        a = np.array([oracle.render(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 5, s, alg, cam, 8)[0].reshape(-1, 3).mean(0) for s in range(6)])
        b = np.array([oracle.render_fp64(sc.vertices, sc.surfaces, sc.lambertian, sc.emissive, 12, 8, 128, 5, 50 + s, alg, cam, 8).reshape(-1, 3).mean(0) for s in range(6)])
        z = (a.mean(0) - b.mean(0)) / np.sqrt(a.var(0, ddof=1) / 6 + b.var(0, ddof=1) / 6)
        assert np.all(np.abs(z) < 4.5), z
