"""Regenerate the golden fixtures under tests/golden/ from the compiled reference.

Run in the development container, where /root/reference exists:

    python tests/golden/make_goldens.py

It builds the reference CPU path (oracle/Makefile, target ``ref``), feeds it GFA inputs
(handcrafted, the reference's own ``example.gfa`` and seeded synthetic pangenomes from
gfaz_amd.synth), and stores what the reference wrote: the traversal fields of each
``.gfaz`` (encoded paths and walks, their lengths, the rule columns after the column prefix
sum, the layer range and delta_round) plus, for the small cases, the ``.gfaz`` bytes and the
text the reference's extract and add-haplotypes commands print. Inputs that the generator
reproduces are stored as a config name and a SHA-256 of the int32 traversal bytes.
"""
from __future__ import annotations

import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from gfaz_py import parse_traversals, read_gfaz  # noqa: E402
from oracle import oracle  # noqa: E402
from gfaz_amd import synth  # noqa: E402

REF = oracle.REF_BIN
EXAMPLE_GFA = "/root/reference/example.gfa"

# The CLI fixture of tests/regression/test_cli_commands.py:24-27, rebuilt from its known
# answers (:114-117, 135, 157-160, 194): the fixture files themselves are not in the
# reference tree.
CLI_FIXTURE = (
    "H\tVN:Z:1.0\n"
    "S\t1\tAC\n"
    "S\t2\tG\n"
    "S\t3\tTTA\n"
    "S\t4\tC\n"
    "L\t1\t+\t2\t+\t0M\n"
    "L\t2\t+\t3\t+\t0M\n"
    "L\t1\t+\t3\t-\t0M\n"
    "L\t3\t-\t4\t+\t0M\n"
    "P\tpathA\t1+,2+,3+\t2M,1M\n"
    "P\tpathB\t1+,3-,4+\t*\n"
    "W\tsampleA\t0\tchr1\t0\t6\t>1>2>3\n"
    "W\tsampleB\t1\tchr2\t*\t*\t>1<3>4\n"
)
APPEND_PATHS = "P\tpathC\t2+,4+,3-\t*\n"
APPEND_WALKS = "W\tsampleC\t0\tchr3\t*\t*\t>2>4>1\n"
APPEND_MIXED = "P\tpathD\t1+,2+\t*\nW\tsampleD\t0\tchr4\t*\t*\t>1>2\n"

# Handcrafted edge cases: single steps, self-loops that give zero deltas, palindromic pairs
# (a, -a), odd and even runs of one repeated pair, reversed copies of forward pairs.
EDGE = (
    "H\tVN:Z:1.0\n"
    + "".join(f"S\t{i}\t{'ACGT'[i % 4] * (1 + i % 3)}\n" for i in range(1, 9))
    + "P\tsingle\t3+\t*\n"
    + "P\tpal\t1+,1-,1+,1-,1+,1-,1+\t*\n"
    + "P\trun_even\t2+,2+,2+,2+,2+,2+\t*\n"
    + "P\trun_odd\t2+,2+,2+,2+,2+,2+,2+\t*\n"
    + "P\tfwd\t1+,2+,3+,4+,5+,6+,7+,8+,1+,2+,3+,4+,5+,6+,7+,8+\t*\n"
    + "P\trev\t8-,7-,6-,5-,4-,3-,2-,1-,8-,7-,6-,5-,4-,3-,2-,1-\t*\n"
    + "P\tmix\t1+,2+,3+,3-,2-,1-,1+,2+,3+,3-,2-,1-\t*\n"
    + "W\ts1\t0\tc1\t*\t*\t>1>2>3>4>1>2>3>4>5\n"
    + "W\ts2\t1\tc2\t0\t9\t<4<3<2<1<4<3<2<1\n"
    + "W\ts3\t0\tc3\t*\t*\t>7\n"
)


def sha(nodes: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(nodes, dtype=np.int32).tobytes()).hexdigest()


def run(*args: str, check: bool = True) -> subprocess.CompletedProcess:
    return subprocess.run([REF, *args], check=check, capture_output=True, text=True)


def compress_case(gfa_path: str, out_gfaz: str, rounds: int, threshold: int, delta: int) -> None:
    run("compress", gfa_path, out_gfaz, str(rounds), str(threshold), str(delta), "4")


def traversal_fields(gfaz_path: str) -> dict:
    g = read_gfaz(gfaz_path)
    rf, rs = g.rules()
    (k, start, end, foff, count), = g.ranges
    return {
        "paths_encoded": g.paths.int32(),
        "paths_lengths": g.sequence_lengths,
        "paths_original_lengths": g.original_path_lengths,
        "walks_encoded": g.walks.int32() if g.walks.original_size else np.zeros(0, np.int32),
        "walks_lengths": g.walk_lengths,
        "walks_original_lengths": g.original_walk_lengths,
        "rules_first": rf,
        "rules_second": rs,
        "range": np.array([k, start, end, foff, count], dtype=np.int64),
        "delta_round": np.array(g.delta_round),
    }


def save(name: str, fields: dict, meta: dict) -> None:
    fields = dict(fields)
    fields["meta"] = np.array(json.dumps(meta, sort_keys=True))
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **fields)
    print(f"wrote {name}.npz", flush=True)


def text_case(name: str, text: str, tmp: str, variants, keep_gfaz: bool = False) -> str:
    gfa = os.path.join(tmp, name + ".gfa")
    with open(gfa, "w") as f:
        f.write(text)
    nodes, lengths, n_paths = parse_traversals(gfa)
    first = None
    for rounds, threshold, delta in variants:
        out = os.path.join(tmp, f"{name}_r{rounds}_t{threshold}_d{delta}.gfaz")
        compress_case(gfa, out, rounds, threshold, delta)
        fields = traversal_fields(out)
        fields.update({"input_nodes": nodes, "input_lengths": lengths})
        save(f"{name}_r{rounds}_t{threshold}_d{delta}", fields,
             {"source": name, "rounds": rounds, "threshold": threshold, "delta_round": delta,
              "n_paths": n_paths, "input_sha256": sha(nodes)})
        if keep_gfaz:
            shutil.copy(out, os.path.join(HERE, f"{name}_r{rounds}_t{threshold}_d{delta}.gfaz"))
        first = first or out
    return first


def synth_case(cfg_name: str, cfg: synth.SynthConfig, tmp: str, variants) -> None:
    tr = synth.generate(cfg)
    gfa = os.path.join(tmp, cfg_name + ".gfa")
    synth.write_gfa(cfg, tr, gfa)
    nodes, lengths, n_paths = parse_traversals(gfa)
    assert np.array_equal(nodes, tr.nodes) and np.array_equal(lengths, tr.lengths)
    for rounds, threshold, delta in variants:
        out = os.path.join(tmp, f"{cfg_name}.gfaz")
        compress_case(gfa, out, rounds, threshold, delta)
        fields = traversal_fields(out)
        meta = {"source": "synth", "config": cfg_name, "rounds": rounds, "threshold": threshold,
                "delta_round": delta, "n_paths": n_paths, "input_sha256": sha(nodes),
                "synth": {k: getattr(cfg, k) for k in cfg.__dataclass_fields__},
                "gfa_bytes": os.path.getsize(gfa), "gfaz_bytes": os.path.getsize(out)}
        save(f"{cfg_name}_r{rounds}_t{threshold}_d{delta}", fields, meta)


def cli_case(tmp: str) -> None:
    """The CLI known answers, reproduced through the reference, plus add-haplotypes output."""
    gfaz = text_case("cli_fixture", CLI_FIXTURE, tmp, [(8, 2, 1)], keep_gfaz=True)
    kat = {}
    kat["extract_path"] = run("extract-path", gfaz, "pathB", "pathA").stdout
    kat["extract_walk"] = run("extract-walk", gfaz, "sampleB", "1", "chr2", "*", "*").stdout
    for label, text in (("paths", APPEND_PATHS), ("walks", APPEND_WALKS), ("mixed", APPEND_MIXED)):
        app = os.path.join(tmp, f"append_{label}.gfa")
        with open(app, "w") as f:
            f.write(text)
        out = os.path.join(tmp, f"cli_fixture_plus_{label}.gfaz")
        res = run("add-haplotypes", gfaz, app, out, check=False)
        kat[f"add_{label}_rc"] = res.returncode
        if res.returncode == 0:
            shutil.copy(out, os.path.join(HERE, f"cli_fixture_plus_{label}.gfaz"))
            if label == "paths":
                kat["after_add_paths"] = run("extract-path", out, "pathA", "pathC").stdout
            else:
                kat["after_add_walks"] = run("extract-walk", out, "sampleC", "0", "chr3", "*",
                                             "*").stdout
    with open(os.path.join(HERE, "cli_fixture_kat.json"), "w") as f:
        json.dump({"append_paths": APPEND_PATHS, "append_walks": APPEND_WALKS,
                   "append_mixed": APPEND_MIXED, "gfa": CLI_FIXTURE, **kat}, f, indent=1)
    print("wrote cli_fixture_kat.json", flush=True)


def add_synth_case(tmp: str) -> None:
    """add-haplotypes of new synthetic P-lines onto the 20k-site pangenome."""
    cfg = synth.CONFIGS["plain_20k"]
    tr = synth.generate(cfg)
    gfa = os.path.join(tmp, "base.gfa")
    synth.write_gfa(cfg, tr, gfa)
    gfaz = os.path.join(tmp, "base.gfaz")
    compress_case(gfa, gfaz, 8, 2, 1)
    new = synth.generate(cfg.scaled(n_haps=6, seed=cfg.seed + 1000))
    off = new.offsets()
    lines = []
    for h in range(new.lengths.size):
        seq = new.nodes[off[h]:off[h + 1]]
        steps = ",".join(f"{abs(int(v))}{'-' if v < 0 else '+'}" for v in seq)
        lines.append(f"P\tnew{h}\t{steps}\t*\n")
    app = os.path.join(tmp, "append_new.gfa")
    with open(app, "w") as f:
        f.writelines(lines)
    out = os.path.join(tmp, "updated.gfaz")
    run("add-haplotypes", gfaz, app, out)
    before = read_gfaz(gfaz)
    after = read_gfaz(out)
    n_old = before.sequence_lengths.size
    enc_after = after.paths.int32()
    old_total = int(before.sequence_lengths.sum(dtype=np.uint64))
    save("plain_20k_add6", {
        "appended_nodes": new.nodes, "appended_lengths": new.lengths,
        "appended_encoded": enc_after[old_total:],
        "appended_encoded_lengths": after.sequence_lengths[n_old:],
        "prefix_unchanged": np.array(np.array_equal(enc_after[:old_total], before.paths.int32())),
    }, {"source": "synth", "config": "plain_20k", "base_rounds": 8, "delta_round": 1,
        "appended_sha256": sha(new.nodes)})


# Step lists the reference parser must read (numeric segment names, so the
# parse_uint32 fast path, src/io/gfa_parser.cpp:1210-1219): empty tokens, a lone
# orientation, the value 0, names with trailing non-digits, a missing orientation,
# leading junk before a W step, and an ID past 2^31.
EDGE_TEXT = (
    "H\tVN:Z:1.1\n"
    "S\t1\tA\n"
    "S\t2\tC\n"
    "S\t3\tG\n"
    "S\t4\tT\n"
    "P\tp1\t1+,,2-,+,0+,3x-,4\t*\n"
    "P\tp2\t12+,2+,1-\t3M,2M\n"
    "P\tp3\t4+,3+,4+,3+\t*\n"
    "W\ts1\t0\tc1\t*\t*\tjunk>1<2>0<3y>4>\n"
    "W\ts2\t1\tc2\t5\t9\t>4>3>4>3<2\n"
    "W\ts3\t2\tc3\t*\t*\t>3000000000<2>1\n"
)
EDGE_TEXT_ERRORS = {
    "no_digits": "H\tVN:Z:1.1\nS\t1\tA\nP\tp\t1+,x+\t*\n",
    "overflow": "H\tVN:Z:1.1\nS\t1\tA\nW\ts\t0\tc\t*\t*\t>1>4294967296\n",
}


def step_lines(gfa_text: str) -> str:
    return "".join(l for l in gfa_text.splitlines(keepends=True) if l[:2] in ("P\t", "W\t"))


def lines_cases(tmp: str) -> None:
    """P/W line text written by the reference's direct writer
    (write_gfa_from_compressed_data), for the step formatter and parser: the
    example and CLI fixtures, a small synthetic pangenome, and edge-case step lists
    (their input lines and the reference's lines after a round trip), plus the
    parser's error messages."""
    def decoded_lines(gfaz: str) -> str:
        out = os.path.join(tmp, "decoded.gfa")
        run("decompress", gfaz, out, "1")
        with open(out) as f:
            return step_lines(f.read())
    for name in ("example_r8_t2_d1", "cli_fixture_r8_t2_d1"):
        with open(os.path.join(HERE, name.split("_r8")[0] + "_lines.txt"), "w") as f:
            f.write(decoded_lines(os.path.join(HERE, name + ".gfaz")))
    cfg = synth.CONFIGS["plain_20k"].scaled(backbone=3000, n_haps=6)
    tr = synth.generate(cfg)
    gfa = os.path.join(tmp, "small.gfa")
    synth.write_gfa(cfg, tr, gfa)
    gz = os.path.join(tmp, "small.gfaz")
    compress_case(gfa, gz, 8, 2, 1)
    with open(os.path.join(HERE, "plain_3k_lines.txt"), "w") as f:
        f.write(decoded_lines(gz))
    edge = os.path.join(tmp, "edge_text.gfa")
    with open(edge, "w") as f:
        f.write(EDGE_TEXT)
    gz = os.path.join(tmp, "edge_text.gfaz")
    compress_case(edge, gz, 8, 2, 1)
    errors = {}
    for label, text in EDGE_TEXT_ERRORS.items():
        bad = os.path.join(tmp, f"bad_{label}.gfa")
        with open(bad, "w") as f:
            f.write(text)
        res = run("compress", bad, os.path.join(tmp, "bad.gfaz"), "8", "2", "1", "1",
                  check=False)
        errors[label] = {"rc": res.returncode, "stderr": res.stderr.strip().splitlines()[-1:]}
    with open(os.path.join(HERE, "edge_text_lines.json"), "w") as f:
        json.dump({"input": step_lines(EDGE_TEXT), "reference_output": decoded_lines(gz),
                   "errors": errors}, f, indent=1)
    print("wrote *_lines.txt, edge_text_lines.json", flush=True)


PAV_BEDS = ("pav_fixture.bed", "pav_path_only_fixture.bed", "pav_walk_fixture.bed")
PANSN_NAMES = ["s0#0#c1", "s0#1#c1", "s1#0#c1", "s1#0#c2:0-100", "s2#0#c1", "s2"]


def analytics_cases(tmp: str) -> None:
    """compute_growth and compute_pav as the reference computes them (through
    oracle/_ref): the CLI fixture with the reference's own BED fixtures, and a small
    synthetic pangenome whose P-lines carry PanSN names, with a BED of windows on one
    of them; every grouping mode."""
    out = {}
    beds_dir = "/root/reference/tests/fixtures"
    for bed in PAV_BEDS:
        shutil.copy(os.path.join(beds_dir, bed), os.path.join(HERE, bed))
    cli = os.path.join(HERE, "cli_fixture_r8_t2_d1.gfaz")
    cfg = synth.CONFIGS["plain_20k"].scaled(backbone=3000, n_haps=6)
    tr = synth.generate(cfg)
    gfa = os.path.join(tmp, "pansn.gfa")
    synth.write_gfa(cfg, tr, gfa)
    with open(gfa) as f:
        lines = f.read().splitlines(keepends=True)
    h = 0
    for i, l in enumerate(lines):
        if l.startswith("P\t"):
            f_ = l.split("\t")
            f_[1] = PANSN_NAMES[h]
            lines[i] = "\t".join(f_)
            h += 1
    with open(gfa, "w") as f:
        f.writelines(lines)
    seg = [len(l.split("\t")[2].rstrip("\n")) for l in lines if l.startswith("S\t")]
    gz = os.path.join(tmp, "pansn.gfaz")
    compress_case(gfa, gz, 8, 2, 1)
    bed = os.path.join(HERE, "pansn_windows.bed")
    with open(bed, "w") as f:
        for k in range(40):
            f.write(f"s0#0#c1\t{k * 997}\t{k * 997 + 1500}\tw{k}\n")
        f.write("s2\t0\t50\n")
        f.write("s1#0#c2:0-100\t100\t100\tempty\n")
    np.savez_compressed(os.path.join(HERE, "analytics_pansn_3k.npz"), nodes=tr.nodes, lengths=tr.lengths,
                        segment_lengths=np.array(seg, np.uint32),
                        names=np.array(json.dumps(PANSN_NAMES)))
    for label, gfaz, beds in (("cli_fixture", cli, PAV_BEDS), ("pansn_3k", gz,
                                                               ("pansn_windows.bed",))):
        case = {"growth": {}, "pav": {}}
        for mode in range(4):
            case["growth"][str(mode)] = json.loads(run("growth", gfaz, str(mode), "2").stdout)
            for b in beds:
                case["pav"][f"{b}/{mode}"] = json.loads(
                    run("pav", gfaz, os.path.join(HERE, b), str(mode), "2").stdout)
        out[label] = case
    with open(os.path.join(HERE, "analytics.json"), "w") as f:
        json.dump(out, f)
    print("wrote analytics.json, analytics_pansn_3k.npz, BED fixtures", flush=True)


def main() -> None:
    oracle.build_reference()
    tmp = tempfile.mkdtemp(prefix="gfaz_goldens_")
    try:
        cli_case(tmp)
        text_case("edge", EDGE, tmp, [(8, 2, 1), (8, 2, 0), (8, 2, 2), (1, 2, 1)], keep_gfaz=True)
        with open(EXAMPLE_GFA) as f:
            example = f.read()
        text_case("example", example, tmp, [(8, 2, 1)], keep_gfaz=True)
        c1 = synth.CONFIGS["plain_20k"]
        synth_case("plain_20k", c1, tmp,
                   [(8, 2, 1), (8, 2, 0), (8, 2, 2), (1, 2, 1), (2, 2, 1), (8, 3, 1)])
        walks = synth.CONFIGS["walks_2m"].scaled(backbone=30_000, n_haps=40, len_max=20_000)
        synth_case("walks_30k", walks, tmp, [(8, 2, 1)])
        add_synth_case(tmp)
        lines_cases(tmp)
        analytics_cases(tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
