"""The algorithmic bytes the bench line's roofline uses.

The canonical figures are the compulsory traffic of the stage-separated RK4
step with 20 stored derivative channels per stage: per site, in channel
accesses, S1 = 36 + 20, S2 = 36 + 20 + 20, S3 = 36 + 16 + 20 + 20,
S4 = 36 + 60 + 36 (356 per step; FP64 x 8 B = 2848 B), and for the mixed
path the same in FP32 plus the FP32 compensation word of y read and written
at stage 4 (1712 B).  A stored stage input u_s replaces the formed reads of
stage s by one 36-channel array and makes stage s-1 also write it.
"""
import bench


def test_canonical_fp64_and_mixed():
    assert bench.STAGE_BYTES["fp64"] == {1: 448, 2: 608, 3: 736, 4: 1056}
    assert sum(bench.STAGE_BYTES["fp64"].values()) == 2848
    assert bench.STAGE_BYTES["mixed"] == {1: 224, 2: 304, 3: 368, 4: 816}
    assert sum(bench.STAGE_BYTES["mixed"].values()) == 1712
    # every formed-input variant moves the canonical bytes
    for v in (0, 2, 10, 11, 311, 320, 325):
        assert bench.stage_bytes_per_site("fp64", v) == bench.STAGE_BYTES["fp64"]


def test_stored_stage_inputs():
    # stages 2-4 stored: S1 y + k1 + u2; S2 u2 + y + k2 + u3; S3 the same; S4 u4 + y + k1..k3 + y'
    full = {1: (36 + 20 + 36) * 8, 2: (36 + 36 + 20 + 36) * 8, 3: (36 + 36 + 20 + 36) * 8,
            4: (36 + 36 + 60 + 36) * 8}
    for v in (315, 324, 332):
        assert bench.stage_bytes_per_site("fp64", v) == full
    assert sum(full.values()) == 4128
    mixed = bench.stage_bytes_per_site("mixed", 333)
    assert mixed[1] == (36 + 20 + 36) * 4
    assert mixed[4] == (36 + 36 + 60 + 36 + 36 + 36) * 4


def test_rebuilt_y():
    # y rebuilt from u2 and k1 at stage 2, from u4 and the k's at stage 4:
    # S2 u2 + k1 + k2 + u3; S4 u4 + k1..k3 + y'
    want = {1: (36 + 20 + 36) * 8, 2: (36 + 20 + 20 + 36) * 8, 3: (36 + 36 + 20 + 36) * 8,
            4: (36 + 60 + 36) * 8}
    assert bench.stage_bytes_per_site("fp64", 334) == want
    assert sum(want.values()) == 3712


def test_measured_traffic_only_at_512_with_source():
    assert bench.measured_traffic("fp64", 332, 4, 256) == (None, None)
    assert bench.measured_traffic("fp64", 999, 4, 512) == (None, None)
    t, src = bench.measured_traffic("fp64", 332, 4, 512)
    assert t == bench.MEASURED_TRAFFIC[("fp64", 332, 4)][0] * 512 ** 3
    assert src.startswith("profiles/") and "commit" in src


def test_default_variants_have_measured_traffic():
    """The automatic variants at 512^3 (FP64 334, mixed 333) carry counter
    figures for all four stages, so the default line's traffic is never null."""
    for prec, v in (("fp64", 334), ("mixed", 333)):
        for s in (1, 2, 3, 4):
            assert (prec, v, s) in bench.MEASURED_TRAFFIC
