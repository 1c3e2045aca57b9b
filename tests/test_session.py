"""Registry, RunSpec command grammar (src/bench/runSpec.ts), seeded noise
(src/mgpu/noise.ts) and state digests (src/mgpu/digest.ts) -- host logic, on
the CPU."""
import math

import numpy as np
import pytest

import oracle as O


def test_presets_and_defaults():
    from tsphere.session import PRESETS, resolve_preset

    assert [p.key for p in PRESETS] == ["schnak-spots", "schnak-coarse", "schnak-fine",
                                        "brussel", "allencahn"]
    _, m, p = resolve_preset("schnak-coarse")
    assert m.key == "schnakenberg" and p == {"a": 0.1, "b": 0.9, "D1": 1e-3, "D2": 2e-2,
                                             "dt": 0.05}
    with pytest.raises(ValueError, match="unknown preset 'nope' \\(have: schnak-spots"):
        resolve_preset("nope")


def test_command_round_trip_and_grid():
    from tsphere.session import config_for_spec, format_command, parse_args

    spec = parse_args(["--preset", "brussel", "--lmax=127", "--steps", "50", "--B", "9.5",
                       "--D1", "0.00001", "--seed", "7"])
    assert spec.lmax == 127 and spec.steps == 50 and spec.params["B"] == 9.5
    assert spec.params["D1"] == 1e-5 and spec.warmup == 100
    cmd = format_command(spec)
    assert cmd == ("python -m tsphere.run --preset brussel --lmax 127 --steps 50 --seed 7 "
                   "--A 3 --B 9.5 --D1 0.00001 --D2 0.0167 --dt 0.02")
    again = parse_args(cmd.split()[3:])
    assert again == spec
    cfg = config_for_spec(spec)
    assert (cfg.nlat, cfg.nphi) == O.grid_for_lmax(127, 3)


@pytest.mark.parametrize("argv,msg", [
    (["lmax"], "unexpected argument 'lmax'"),
    (["--lmax"], "--lmax needs a value"),
    (["--lmax", "0"], "--lmax must be an integer >= 1 (got '0')"),
    (["--steps", "1.5"], "--steps must be an integer >= 1 (got '1.5')"),
    (["--seed", "x"], "--seed must be a number (got 'x')"),
    (["--a", "inf"], "--a must be a number (got 'inf')"),
    (["--eps2", "1"], "unknown option(s): --eps2\nparameters of Schnakenberg: --a --b --D1 "
                      "--D2 --dt"),
])
def test_parse_errors(argv, msg):
    from tsphere.session import parse_args

    with pytest.raises(ValueError) as e:
        parse_args(argv)
    assert str(e.value) == msg


def test_seeded_noise_matches_oracle_randn():
    """mulberry32 + Box-Muller in pairs (noise.ts:27-51), against the
    oracle's C mulberry32 stream with the same Box-Muller."""
    from tsphere.session import make_rand, seeded_noise

    r = make_rand(123)
    u = np.array([r() for _ in range(1000)])
    assert np.array_equal(u, O.mulberry32(123, 1000))
    n = seeded_noise(1001, 1e-2, 5)
    ref = 1e-2 * O.randn(5, 1001)
    assert np.max(np.abs(n - ref.astype(np.float32))) <= 1e-9


def test_digest():
    from tsphere.session import digest_drift, digest_of, format_digest, rel_l2

    v = np.array([1.0, -2.0, 3.5, 0.25], np.float32)
    d = digest_of(v, "fft", "x")
    assert (d.n, d.min, d.max) == (4, -2.0, 3.5)
    assert d.mean == pytest.approx(0.6875) and d.rms == pytest.approx(math.sqrt(17.3125 / 4))
    assert format_digest(d) == ("n=4 min=-2.00000000 max=3.50000000 mean=0.687500000 "
                                "rms=2.08041462 fourier=fft")
    d2 = digest_of(v * 1.001, "fft", "y")
    assert digest_drift(d, d2) == pytest.approx(0.001 / 1.001, rel=1e-3)
    # toPrecision(9) formatting of the JS engine, at the edges of fixed notation
    from tsphere.session import _prec9
    assert [_prec9(x) for x in (1.23e-7, 123456789012.0, 0.0)] == [
        "1.23000000e-7", "1.23456789e+11", "0.00000000"]
    assert rel_l2(v.astype(np.float64) * 1.001, v) == pytest.approx(0.001, rel=1e-9)
