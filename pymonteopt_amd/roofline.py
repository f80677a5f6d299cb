"""Work model of the walk, for roofline fractions.

The unit of work is one packet. Its VALU work is counted from the walk's own event
counters (interactions, face crossings, Fresnel crossings, roulette tests) times a fixed
per-event cost in VALU lane-op issue slots, counted from the reference algorithm
(pymopt/voxel/_kernel.py:61-165): transcendental ops (log, sin, cos, sqrt, rcp, rsqrt)
count 4 slots (quarter rate), acos/asin/tan count 16 (software sequences); an interaction
includes its Philox block (10 rounds). The table is frozen: every reported fraction uses
these weights, whatever the kernel actually issues.
"""
from __future__ import annotations

# lane-op slots per event
WEIGHTS = {
    "interaction": 184.0,   # hop end + drop + roulette test + HG spin + rotation + renorm + 4 uniforms
    "crossing": 45.0,       # plain crossing: boundary search + move + neighbour label fetch
    "fresnel": 115.0,       # crossing with an index change (adds Fresnel, refraction, 1 uniform)
    "roulette": 30.0,       # roulette test taken (w <= wth)
}

# MI355X (gfx950): 256 CUs x 4 SIMD x 16 lanes per clock at 2.4 GHz peak engine clock
N_CU = 256
LANES_PER_CU_CLK = 64
CLOCK_HZ = 2.4e9
PEAK_VALU_LANE_OPS = N_CU * LANES_PER_CU_CLK * CLOCK_HZ     # 3.93e13 lane-op/s, one op per lane per clock
# AMD's published MI355X FP32 vector peak, 157.3 TFLOP/s, counts an FMA as 2 flops and
# assumes packed (two-wide) FP32; in the lane-op slots of WEIGHTS it is 7.86e13 slots/s
PEAK_FP32_PUBLISHED_FLOPS = 157.3e12
PEAK_VALU_PUBLISHED_LANE_OPS = PEAK_FP32_PUBLISHED_FLOPS / 2
PEAK_HBM_BPS = 8.0e12


def ops_per_packet(events: dict) -> float:
    n = max(1, events["packets"])
    plain_cross = events["crossings"] - events["fresnel"]
    ops = (events["interactions"] * WEIGHTS["interaction"] + plain_cross * WEIGHTS["crossing"]
           + events["fresnel"] * WEIGHTS["fresnel"] + events["roulette"] * WEIGHTS["roulette"])
    return ops / n


def algorithmic_bytes_per_packet(events: dict, absorbed: bool = True) -> float:
    """HBM-side bytes the algorithm needs per packet: one label byte per face crossing and,
    with the absorbed tally, one 8-byte read-modify-write (16 B) per interaction."""
    n = max(1, events["packets"])
    b = events["crossings"] * 1.0
    if absorbed:
        b += events["interactions"] * 16.0
    return b / n
