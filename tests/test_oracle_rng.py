"""The oracle's MCG59 leap-frog stream against Intel MKL golden vectors.

Fixture tests/golden/mcg59_mkl.json was written by tests/golden/gen_mcg59_fixture.c
calling MKL 2021.4.0 VSL (vslNewStream(MCG59, 777), vslLeapfrogStream(k, 1e6),
vdRngUniform STD) -- the calls of the reference's newdel_exp.cpp:244-296.
"""
import json
import os
import struct

import numpy as np

FIX = os.path.join(os.path.dirname(__file__), "golden", "mcg59_mkl.json")


def _bits(v):
    return "%016x" % struct.unpack("<Q", struct.pack("<d", v))[0]


def test_leapfrog_streams_bit_exact(oracle_mod):
    fx = json.load(open(FIX))
    for s in fx["streams"]:
        u = oracle_mod.mcg59_uniform(s["k"], fx["mns"], len(s["bits"]))
        assert [_bits(x) for x in u] == s["bits"], "stream k=%d" % s["k"]


def test_plain_stream_bit_exact(oracle_mod):
    fx = json.load(open(FIX))
    # nstreams = 1 leap-frog is the plain sequence
    u = oracle_mod.mcg59_uniform(0, 1, len(fx["plain_4096"]))
    assert [_bits(x) for x in u] == fx["plain_4096"]


def test_first_output_is_the_seed():
    # x_0 = 777, so the first draw of stream 0 is 777 / 2^59 (why seed 0 gives
    # trajectory 0 an almost-zero first eta).
    fx = json.load(open(FIX))
    assert fx["streams"][0]["bits"][0] == _bits(777.0 / 2.0 ** 59)


def test_uniform_range(oracle_mod):
    u = oracle_mod.mcg59_uniform(5, 1000000, 10000)
    assert np.all(u >= 0.0) and np.all(u < 1.0)
    assert abs(u.mean() - 0.5) < 0.02
