"""Loading of the committed golden fixtures (written by tests/golden/make_goldens.py from
the compiled reference CPU path)."""
from __future__ import annotations

import glob
import hashlib
import json
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names(pattern: str = "*") -> list[str]:
    out = []
    for p in sorted(glob.glob(os.path.join(GOLDEN_DIR, pattern + ".npz"))):
        n = os.path.basename(p)[:-4]
        if not n.endswith("_add6") and not n.startswith("analytics_"):
            out.append(n)
    return out


def load(name: str) -> dict:
    with np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


def inputs(g: dict):
    """(nodes, lengths, n_paths) of a golden case; synthetic inputs are regenerated and
    checked against the stored SHA-256."""
    meta = g["meta"]
    if "input_nodes" in g:
        nodes, lengths = g["input_nodes"], g["input_lengths"]
    else:
        from gfaz_amd import synth
        cfg = synth.SynthConfig(**meta["synth"])
        tr = synth.generate(cfg)
        nodes, lengths = tr.nodes, tr.lengths
    digest = hashlib.sha256(np.ascontiguousarray(nodes, dtype=np.int32).tobytes()).hexdigest()
    assert digest == meta["input_sha256"], "regenerated input differs from the golden's input"
    return nodes, lengths, int(meta["n_paths"])


def expected(g: dict) -> dict:
    """Traversal grammar as one flat stream (paths then walks)."""
    return {
        "encoded": np.concatenate([g["paths_encoded"], g["walks_encoded"]]).astype(np.int32),
        "encoded_lengths": np.concatenate([g["paths_lengths"], g["walks_lengths"]]).astype(np.uint32),
        "original_lengths": np.concatenate([g["paths_original_lengths"],
                                            g["walks_original_lengths"]]).astype(np.uint32),
        "rules_first": g["rules_first"].astype(np.int32),
        "rules_second": g["rules_second"].astype(np.int32),
        "start_id": int(g["range"][1]),
        "end_id": int(g["range"][2]),
        "delta_round": int(g["delta_round"]),
    }
