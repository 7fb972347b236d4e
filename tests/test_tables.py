"""Tables the backend and the oracle restate, pinned against the reference's
own sources (oracle/_ref/ref_tables, built by oracle/Makefile from
util/promote.hpp, util/{DataType,VariantType,ReductionType,BinaryType}.cpp
and runtime/Pattern.cpp)."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "ref_tables")


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    return json.loads(subprocess.run([REF], check=True, capture_output=True, text=True).stdout)


TYPES = list(range(1, 12))


def test_oracle_promote_matches_reference(ref):
    from oracle import ref as O
    for a in TYPES:
        for b in TYPES:
            assert O.promote(a, b) == ref["promote"][a - 1][b - 1], (a, b)


def _const(M, dt):
    v = M.Variant(1, dt) if dt != M.B8 else M.Variant(True, dt)
    return M.Raster(M._lib.ma_constant(v, M.Array(), dt, M.ROW + M.BLK, M.Array()))


def test_backend_promote_matches_reference(M, ref):
    for a in TYPES:
        for b in TYPES:
            r = _const(M, a) + _const(M, b)
            assert r.datatype() == ref["promote"][a - 1][b - 1], (a, b)


def test_backend_fusion_rules_match_reference(M, ref):
    for top, bot, pipe, flat in ref["fusion"]:
        got = M._lib.ma_fusionRules(top, bot)
        if got < 0:  # the reference asserts for this pair
            continue
        assert (got & 1) == pipe, ("pipe", top, bot)
        assert ((got >> 1) & 1) == flat, ("flat", top, bot)


def test_mask_literals_match_reference(M, ref):
    """Float mask weights are quantised as the reference prints them (%f), and
    integer masks are still printed into the kernel."""
    from oracle import ref as O
    for f, s in ref["float_literal"].items():
        want = np.float32(float(s))
        assert np.float32(O._quantize(float(np.float32(float(f))))) == want
        r = M.from_array(np.zeros((8, 8), np.float32), (8, 8))
        c = M.convolve(r, [[float(np.float32(float(f))), 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
        g = M.plan(c)["groups"][-1]
        # floating weights reach the kernel as arguments, quantised as the reference prints them
        assert re.search(r"\+= W\.w\d+\[H1 \+ \(-1\)\]\[PAD \+ v \+ \(-1\)\] \* a\.m\d+\[0\];", g["source"])
        assert np.float32(g["mask_args"][0][0]) == want, (f, g["mask_args"][0][0])
    r = M.from_array(np.zeros((8, 8), np.int32), (8, 8))
    src = M.plan(M.convolve(r, [[3, 0, -2], [0, 1, 0], [0, 0, 7]]))["groups"][-1]["source"]
    assert re.search(r"\+= W\.w\d+\[H1 \+ \(-1\)\]\[PAD \+ v \+ \(1\)\] \* \(?-2\)?;", src), src


def test_binary_spellings_match_reference(ref):
    from map_amd import map as M
    for code, sym in ref["binary_code"].items():
        assert M.BinaryTypeId[int(code)] in M.BinaryTypeId
    # the operators the code generator prints (C spellings) agree with the reference
    ops = {1: "+", 2: "-", 3: "*", 4: "/", 6: "==", 7: "!=", 8: "<", 9: ">", 10: "<=", 11: ">=", 12: "&&", 13: "||",
           14: "&", 15: "|", 16: "^", 17: "<<", 18: ">>"}
    for k, v in ops.items():
        assert ref["binary_code"][str(k)] == v
