"""Inputs of the growth/PAV goldens (tests/golden/analytics.json): decoded traversals,
names and W-line fields, segment lengths, BED ranges."""
import json
import os

import numpy as np

import goldens
from gfaz_amd import analytics as an

GD = goldens.GOLDEN_DIR


def case(label):
    if label == "cli_fixture":
        g = goldens.load("cli_fixture_r8_t2_d1")
        nodes, lengths, n_paths = goldens.inputs(g)
        fx = json.load(open(os.path.join(GD, "cli_fixture_kat.json")))
        names, walks, seg = [], [], []
        for line in fx["gfa"].splitlines():
            f = line.split("\t")
            if f[0] == "S":
                seg.append(len(f[2]))
            elif f[0] == "P":
                names.append(f[1])
            elif f[0] == "W":
                walks.append((f[1], int(f[2]), f[3], -1 if f[4] == "*" else int(f[4]),
                              -1 if f[5] == "*" else int(f[5])))
        beds = ("pav_fixture.bed", "pav_path_only_fixture.bed", "pav_walk_fixture.bed")
    else:
        z = np.load(os.path.join(GD, "analytics_pansn_3k.npz"), allow_pickle=False)
        nodes, lengths, seg = z["nodes"], z["lengths"], z["segment_lengths"].tolist()
        names, walks = json.loads(str(z["names"])), []
        beds = ("pansn_windows.bed",)
    ref = json.load(open(os.path.join(GD, "analytics.json")))[label]
    return {"nodes": np.asarray(nodes, np.int32), "lengths": np.asarray(lengths, np.uint32),
            "names": names, "walks": walks, "seg": np.array(seg, np.uint32), "beds": beds,
            "ref": ref}


def groups(c, mode, kind):
    if kind == "growth":
        keys = an.growth_keys(c["names"], [w[:3] for w in c["walks"]], mode)
        if keys is None:
            n = c["lengths"].size
            return np.arange(n, dtype=np.uint32), n
        g, names = an.group_ids(keys)
        return g, len(names)
    g, names = an.group_ids(an.pav_keys(c["names"], c["walks"], mode))
    return g, names
