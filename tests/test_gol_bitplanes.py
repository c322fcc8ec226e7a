"""The bit-sliced B3/S23 update of the packed GameOfLife kernels, checked on the CPU.

The kernels evaluate the rule on 32 cells per word with three-input `v_bitop3_b32`
operations, each given by an 8-bit truth table (bit a*4 + b*2 + c of the immediate is
the result for input bits a, b, c).  This test reads the immediates and the operand order
from the kernel source itself (xor3, maj and life_rule in gol_bits.hip), evaluates them,
and compares the result with the rule's definition, S == 3 or (alive and S == 4) for the
3x3 sum S including the centre (the rule of GameOfLife/src/game_appfis.hpp:121-135), over every
combination of three 2-bit row sums and centre state.
"""
import itertools
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "appfis_amd", "csrc", "gol_bits.hip")


def _body(src, name):
    m = re.search(r"uint32_t\s+" + name + r"\s*\(([^)]*)\)\s*\{(.*?)\n\}", src, re.S)
    assert m, name
    params = [p.split()[-1] for p in m.group(1).split(",")]
    return params, m.group(2)


_CALL = re.compile(r"__builtin_amdgcn_bitop3_b32\(\s*(\w+),\s*(\w+),\s*(\w+),\s*(0x[0-9A-Fa-f]+)\s*\)")


def _bitop3(a, b, c, imm):
    return (imm >> (a * 4 + b * 2 + c)) & 1


def _compile(src, name):
    """Returns f(**params) for a function made of bitop3 assignments and a return."""
    params, body = _body(src, name)
    steps = []
    for line in body.strip().splitlines():
        line = line.strip()
        if not line or line.startswith("//"):
            continue
        call = _CALL.search(line)
        assert call, f"{name}: unexpected line {line!r}"
        m = re.match(r"(?:const\s+uint32_t\s+(\w+)\s*=|return)", line)
        assert m, line
        steps.append((m.group(1), call.group(1), call.group(2), call.group(3), int(call.group(4), 16)))

    def f(**env):
        assert sorted(env) == sorted(params)
        env = dict(env)
        for dst, a, b, c, imm in steps:
            v = _bitop3(env[a], env[b], env[c], imm)
            if dst is None:
                return v
            env[dst] = v
        raise AssertionError(f"{name} has no return")

    return params, f


def test_life_rule_matches_b3s23_for_every_neighbourhood():
    src = open(SRC).read()
    _, xor3 = _compile(src, "xor3")
    _, maj = _compile(src, "maj")
    params, rule = _compile(src, "life_rule")
    assert params == ["s0", "c0", "v1", "c1", "alive"]
    for a, b, c in itertools.product(range(2), repeat=3):
        assert xor3(a=a, b=b, c=c) == (a + b + c) & 1
        assert maj(a=a, b=b, c=c) == int(a + b + c >= 2)
    checked = 0
    # each row contributes its horizontal 3-cell sum h = h0 + 2*h1 (0..3); the centre
    # row's sum includes the cell itself, so alive = 1 needs that row's sum >= 1
    for (ha, hb, hc), alive in itertools.product(itertools.product(range(4), repeat=3), range(2)):
        if alive and hb == 0:
            continue
        s0 = xor3(a=ha & 1, b=hb & 1, c=hc & 1)
        c0 = maj(a=ha & 1, b=hb & 1, c=hc & 1)
        v1 = xor3(a=ha >> 1, b=hb >> 1, c=hc >> 1)
        c1 = maj(a=ha >> 1, b=hb >> 1, c=hc >> 1)
        S = ha + hb + hc
        want = int(S == 3 or (alive and S == 4))
        assert rule(s0=s0, c0=c0, v1=v1, c1=c1, alive=alive) == want, (ha, hb, hc, alive)
        checked += 1
    assert checked == 64 + 48


def test_planted_rule_error_is_caught():
    # the same rule with the survival term (S == 4, live centre) dropped must disagree
    src = open(SRC).read()
    bad = src.replace("alive, s0, q, 0x20", "alive, s0, q, 0x00")
    assert bad != src
    _, rule = _compile(bad, "life_rule")
    _, xor3 = _compile(src, "xor3")
    _, maj = _compile(src, "maj")
    ha, hb, hc = 1, 2, 1  # S = 4 with a live centre: survives under B3/S23
    s0 = xor3(a=1, b=0, c=1)
    c0 = maj(a=1, b=0, c=1)
    v1 = xor3(a=0, b=1, c=0)
    c1 = maj(a=0, b=1, c=0)
    assert ha + hb + hc == 4 and rule(s0=s0, c0=c0, v1=v1, c1=c1, alive=1) == 0
