"""Markdown summary of the per-frame parity statistics the GPU parity tests append to $HUSH_PARITY_LOG.

usage: HUSH_PARITY_LOG=gpurun_out/parity.jsonl python -m pytest tests/test_gpu_parity.py tests/test_gpu_fullsize.py \
           tests/test_gpu_api.py -m gpu
       python tests/parity_summary.py gpurun_out/parity.jsonl > profiles/parity_summary.md

Teacher-forced cases ("tf-*"): every pass runs on the oracle's copy of the GPU state, so each row is one pass's
own deviation; per (pass, surface) the table gives the smallest exact-match fraction over the frames, the largest
difference in binary16 ulps of the oracle's value and the largest relative error.  End-to-end cases ("e2e-*"):
the GPU and the oracle run the sequence independently; one row per frame.
"""
from __future__ import annotations

import collections
import json
import sys


def load(path):
    cases = collections.OrderedDict()
    with open(path) as fh:
        for line in fh:
            rec = json.loads(line)
            cases[rec["case"]] = rec["stats"]          # a rerun of a case replaces the earlier record
    return cases


def fmt(x, digits=3):
    if isinstance(x, float):
        if x == 0.0:
            return "0"
        if abs(x) >= 1e-2 and abs(x) < 1e4:
            return f"{x:.{digits + 1}g}"
        return f"{x:.{digits}g}"
    return str(x)


def teacher_forced(case, stats):
    rows = collections.OrderedDict()
    for s in stats:
        key = (s["pass_"], s["surface"])
        r = rows.setdefault(key, {"frames": 0, "n": 0, "exact": 1.0, "max_ulp": 0.0, "max_rel": 0.0})
        r["frames"] += 1
        r["n"] += s.get("n", 0)
        r["exact"] = min(r["exact"], s["exact"])
        if s["surface"] == "mask":
            r["max_ulp"] = max(r["max_ulp"], 0.0 if s["exact"] == 1.0 else float("nan"))
        else:
            r["max_ulp"] = max(r["max_ulp"], s["max_ulp"])
            r["max_rel"] = max(r["max_rel"], s["max_rel"])
    out = [f"### {case}", "", "| pass | surface | frames | values | min exact fraction | max ulp | max rel |",
           "|---|---|---|---|---|---|---|"]
    for (p, surf), r in rows.items():
        out.append(f"| {p} | {surf} | {r['frames']} | {r['n']} | {r['exact']:.6f} | {fmt(r['max_ulp'])} | "
                   f"{fmt(r['max_rel'])} |")
    return out + [""]


def end_to_end(case, stats):
    out = [f"### {case}", "", "| frame | exact fraction | values beyond 1e-3 rel | beyond 1e-2 rel | max rel | "
           "99.99% ulp | max ulp | validity masks equal |", "|---|---|---|---|---|---|---|---|"]
    for s in stats:
        out.append(f"| {s.get('frame', '')} | {s['exact']:.6f} | {fmt(s['frac_rel_gt_1e-3'])} | "
                   f"{fmt(s.get('frac_rel_gt_1e-2', 0.0))} | {fmt(s['max_rel'])} | {fmt(s['p9999_ulp'])} | "
                   f"{fmt(s['max_ulp'])} | {s.get('mask_exact', '')} |")
    worst = max(stats, key=lambda s: s["frac_rel_gt_1e-3"])
    out += ["", f"Largest fraction beyond 1e-3 over the frames: {fmt(worst['frac_rel_gt_1e-3'])} (frame "
            f"{worst.get('frame', '')}); largest relative error {fmt(max(s['max_rel'] for s in stats))}.", ""]
    return out


def main(argv):
    cases = load(argv[1])
    lines = ["# Parity of the HIP path against the oracle", "",
             "Produced by `HUSH_PARITY_LOG=gpurun_out/parity.jsonl python -m pytest tests/test_gpu_parity.py "
             "tests/test_gpu_fullsize.py tests/test_gpu_api.py -m gpu` on one MI355X, then "
             "`python tests/parity_summary.py gpurun_out/parity.jsonl`.", "",
             "Values are compared as float64; ulps are binary16 ulps of the oracle's value (2^-24 below the normal "
             "range); relative errors are |gpu - oracle| / max(|oracle|, 1e-3).  `ieee` cases use the IEEE math "
             "policy (the oracle's operation-exact mirror), `fast` cases the hardware reciprocal / sqrt / exp2 / "
             "log2 in the spatial filters.", "",
             "Residual.  Teacher-forced, every pass of the IEEE policy matches the oracle on at least 99.99% of "
             "its values and the rest differ by one binary16 ulp (relative 2^-10 = 9.8e-4 at most): the fp32 "
             "results agree to within an fp32 ulp (the GPU's exp / log / pow against the host's), and the "
             "round-toward-zero binary16 store (hush.js:98) turns that into one binary16 ulp when the value lies "
             "that close to a binary16 boundary.  End to end, these differences ride the history (the temporal "
             "recurrence and the TAA), so a few tenths of a percent of the values differ, 99.99% of them by at "
             "most one ulp, and at most about 4e-5 of the values lie beyond 1e-3 relative.  The isolated larger "
             "relative errors of the long 1920x1080 sequence come from two amplifiers of a one-ulp input "
             "difference: the TAA's inverse tone map 1 / (1 - lum) near its clamp (1e-4) and the recombination of "
             "levels, where nearly equal terms cancel; teacher-forced, both passes stay within one ulp.", ""]
    tf = [(c, s) for c, s in cases.items() if c.startswith("tf-")]
    e2e = [(c, s) for c, s in cases.items() if c.startswith("e2e-")]
    other = [(c, s) for c, s in cases.items() if not c.startswith(("tf-", "e2e-"))]
    if tf:
        lines += ["## Teacher-forced passes", ""]
        for c, s in tf:
            lines += teacher_forced(c, s)
    if e2e:
        lines += ["## End to end", ""]
        for c, s in e2e:
            lines += end_to_end(c, s)
    if other:
        lines += ["## Other sequences", ""]
        for c, s in other:
            lines += end_to_end(c, s)
    sys.stdout.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main(sys.argv)
