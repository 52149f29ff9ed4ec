"""Writes known_answers.json: known-answer vectors for the oracle, transcribed from the
reference's own tests and from the published algorithms (values only; no reference code).

Sources:
  * Source/Tools/FalcorTest/Tests/Utils/BitTricksTests.cpp:57-58 (bit interleave examples)
  * Source/Tools/FalcorTest/Tests/DiffRendering/Material/DiffMaterialTests.cpp:86-90,110-114
    (PBRTDiffuse backward derivative: config and expected gradients, tolerance 1e-3)
  * Source/Tools/FalcorTest/Tests/DiffRendering/SceneGradientsTest.cpp:36-72 (hash-grid
    aggregation: 1024 atomic adds of 10^i, rel. tolerance 1e-6, hash size 64)
  * SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c) from state 0: first three outputs
  * xoshiro128** (Blackman & Vigna) from state {1, 2, 3, 4}: first two outputs, derived by hand:
    rotl(1*5, 7)*9 = 5760; after one step s0 = 1 ^ (4 ^ 2) = 7, rotl(7*5, 7)*9 = 40320.
"""
import json
from pathlib import Path

DATA = {
    "interleave": [
        {"x": 0xE38E, "y": 0xBE8B, "bits": 16, "expect": 0xDEADC0DE},
        {"x": 0xE38E & 0xFFF, "y": 0xBE8B & 0xFFF, "bits": 12, "expect": 0x00ADC0DE},
    ],
    "splitmix64_state0": [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F],
    "xoshiro128ss_state1234": [5760, 40320],
    "diff_pbrt_diffuse": {
        "wi": [0.3, 0.2, 0.8], "wo": [-0.1, -0.3, 0.9], "base_color": [0.9, 0.6, 0.2],
        "grad_albedo": [0.3003115, 0.3003115, 0.3003115], "grad_wo_z": 0.5411268, "tol": 1e-3,
    },
    "aggregate_gradients": {"dim": 3, "adds": 1024, "hash_size": 64, "rel_tol": 1e-6},
}

if __name__ == "__main__":
    out = Path(__file__).with_name("known_answers.json")
    out.write_text(json.dumps(DATA, indent=1))
    print(f"wrote {out}")
