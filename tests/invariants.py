"""Size-independent properties of a traversal grammar (the reference's algorithm fixes
them: src/workflows/compression_workflow.cpp:548-700, src/grammar/rule_generator.cpp:
148-260), checked with whole-array numpy so they run on billions of symbols."""
import numpy as np


def check_grammar(encoded, rules_first, rules_second, start_id, end_id, rules_per_round):
    """Every rule key is canonical (signed min of the packed key and its reverse); the
    rules of each round are sorted ascending by signed key; a rule's children are
    earlier rules or node IDs and its depth equals the round that made it; every
    encoded symbol is a node ID below start_id or a rule ID below end_id."""
    f = np.asarray(rules_first, dtype=np.int64)
    s = np.asarray(rules_second, dtype=np.int64)
    assert f.size == s.size == end_id - start_id == sum(rules_per_round)
    key = (f << 32) | (s & 0xFFFFFFFF)
    rev = ((-s) << 32) | ((-f) & 0xFFFFFFFF)
    assert np.all(key <= rev), "every rule key is canonical"
    depth = np.zeros(f.size, dtype=np.int16)
    o = 0
    for rnd, cnt in enumerate(rules_per_round, start=1):
        assert np.all(np.diff(key[o:o + cnt]) > 0), f"round {rnd}: rules sorted by signed key"
        ids = start_id + np.arange(o, o + cnt, dtype=np.int64)
        d = np.zeros(cnt, dtype=np.int16)
        for col in (f, s):
            ch = np.abs(col[o:o + cnt])
            is_rule = (ch >= start_id) & (ch < end_id)
            assert np.all(~is_rule | (ch < ids)), f"round {rnd}: children precede the rule"
            cd = np.where(is_rule, depth[np.where(is_rule, ch - start_id, 0)], 0)
            d = np.maximum(d, cd)
        d += 1
        assert np.all(d == rnd), f"round {rnd}: a rule's depth equals its round"
        depth[o:o + cnt] = d
        o += cnt
    enc = np.abs(np.asarray(encoded, dtype=np.int64))
    assert np.all(enc < end_id), "encoded symbols are node IDs or rule IDs"
