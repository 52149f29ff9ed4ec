"""Writes known_answers.json: known-answer vectors for the oracle, transcribed from the
reference's own tests and from the published algorithms (values only; no reference code).

Sources:
  * Source/Tools/FalcorTest/Tests/Utils/BitTricksTests.cpp:57-58 (bit interleave examples)
  * Source/Tools/FalcorTest/Tests/DiffRendering/Material/DiffMaterialTests.cpp:86-90,110-114
    (PBRTDiffuse backward derivative: config and expected gradients, tolerance 1e-3)
  * Source/Tools/FalcorTest/Tests/DiffRendering/SceneGradientsTest.cpp:36-72 (hash-grid
    aggregation: 1024 atomic adds of 10^i, rel. tolerance 1e-6, hash size 64)
  * SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c) from state 0: first three outputs
  * xoshiro128** as the reference states it (output from state[0]; Xoshiro.slang:54-69 and the CPU
    reference struct of PseudorandomTests.cpp:42-63) from state {1, 2, 3, 4}: eight outputs. The
    first three by hand (state after each step in hex, s0 s1 s2 s3):
      out0 = rotl(1*5, 7)*9 = 5760;       t = 2<<9 = 0x400 -> 7 0 0x402 0x3000 (rotl(4^2, 11))
      out1 = rotl(7*5, 7)*9 = 40320;      t = 0           -> 0x3007 0x405 0x405 0x1800000
      out2 = rotl(0x3007*5, 7)*9 = 70819200 (0xF023 << 7 = 0x781180, * 9)
    out2 is the first output that depends on rotl(s3, 11) and out4 the first through s2 ^= s1 << 9;
    outputs 3-7 from the same eight state operations in exact 32-bit integer arithmetic below.
  * jenkinsHash (HashUtils.slang:39-48 and the CPU function of HashUtilsTests.cpp:42-51) of nine
    inputs. jenkinsHash(0) by hand, the value after each of the six steps:
      0x7ed55d16, 0xb9b490f0 (^ 0xc761c23c ^ 0xfda), 0x069d16a1, 0xe012390d, 0xde14483a, 0x6b4ed927
    = 1800329511; the other inputs from the same six steps in exact 32-bit arithmetic below.
"""
import json
from pathlib import Path

M32 = 0xFFFFFFFF


def _rotl(x, k):
    return ((x << k) | (x >> (32 - k))) & M32


def _xoshiro128ss(state, n):
    s = list(state)
    out = []
    for _ in range(n):
        out.append((_rotl((s[0] * 5) & M32, 7) * 9) & M32)
        t = (s[1] << 9) & M32
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 11)
    return out


def _jenkins(a):
    a = ((a + 0x7ED55D16) + (a << 12)) & M32
    a = ((a ^ 0xC761C23C) ^ (a >> 19)) & M32
    a = ((a + 0x165667B1) + (a << 5)) & M32
    a = ((a + 0xD3A2646C) ^ (a << 9)) & M32
    a = ((a + 0xFD7046C5) + (a << 3)) & M32
    a = ((a ^ 0xB55A4F09) ^ (a >> 16)) & M32
    return a


JENKINS_INPUTS = [0, 1, 2, 3, 255, 65535, 1 << 20, 0xDEADBEEF, 0xFFFFFFFF]

DATA = {
    "interleave": [
        {"x": 0xE38E, "y": 0xBE8B, "bits": 16, "expect": 0xDEADC0DE},
        {"x": 0xE38E & 0xFFF, "y": 0xBE8B & 0xFFF, "bits": 12, "expect": 0x00ADC0DE},
    ],
    "splitmix64_state0": [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F],
    "xoshiro128ss_state1234": _xoshiro128ss([1, 2, 3, 4], 8),
    "jenkins": {"inputs": JENKINS_INPUTS, "hashes": [_jenkins(x) for x in JENKINS_INPUTS]},
    "diff_pbrt_diffuse": {
        "wi": [0.3, 0.2, 0.8], "wo": [-0.1, -0.3, 0.9], "base_color": [0.9, 0.6, 0.2],
        "grad_albedo": [0.3003115, 0.3003115, 0.3003115], "grad_wo_z": 0.5411268, "tol": 1e-3,
    },
    "aggregate_gradients": {"dim": 3, "adds": 1024, "hash_size": 64, "rel_tol": 1e-6},
}

assert DATA["xoshiro128ss_state1234"][:3] == [5760, 40320, 70819200]  # the hand derivation above
assert DATA["jenkins"]["hashes"][0] == 0x6B4ED927 == 1800329511

if __name__ == "__main__":
    out = Path(__file__).with_name("known_answers.json")
    out.write_text(json.dumps(DATA, indent=1))
    print(f"wrote {out}")
