"""Writes tests/golden/resparse_search_logs.json from the reference's ReSparse search logs.

The reference keeps the output of every keep-ratio search it ran
(data/matrices_for_testing/*/32-cores/results/*_search_resparse_<threshold>_CK<ck>_split<split>.txt,
written by datasets/search_best_resparse_ratio.py:55-116 through quantify_util.py:413-437 for the
embedding matrices, and by self_generated_data/search_best_resparse_ratio.py:45-100 through
self_generated_data/csr_util.py:215-262 for the generated ones).  Each epoch prints the ratio it
tried ("Resparse ... with top_percent=..."), the entry counts before and after pruning, the mean
precision@K for K = 1, 8, 16, 32, 50, 75, 100 (rounded to 0.01 % in the first form, unrounded in the
second) and the convergence so far; the last epoch prints the ratio the search settled on.  This
script keeps those numbers (data only) so that the tests can replay the search and the metric:

    python tests/golden/make_resparse_logs.py [reference root, default /root/reference]
"""
import glob
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path):
    name = os.path.basename(path)
    m = re.search(r"_search_resparse_([0-9.]+)_CK(\d+)_split(\d+)\.txt$", name)
    thr, ck, split = float(m.group(1)), int(m.group(2)), int(m.group(3))
    epochs, conv_lists, converged = [], [], None
    cur = None
    for line in open(path, encoding="utf-8", errors="replace"):
        line = line.strip()
        if line.startswith("Epoch:"):
            cur = {"epoch": int(line.split()[1]), "precision_pct": []}
            epochs.append(cur)
        elif re.match(r"Resparse (SpM|CSR) with top_percent=", line):
            cur["top_percent"] = float(line.split("=")[1].rstrip("%"))
        elif re.match(r"Before Re[sS]parse:", line):
            cur["nnz_before"] = int(re.search(r"entries: (\d+)", line).group(1))
        elif re.match(r"After Re[sS]parse:", line):
            cur["nnz_after"] = int(re.search(r"entries: (\d+)", line).group(1))
        elif line.startswith("K=") and re.search(r"avg_precis?i?on:", line):
            cur["precision_pct"].append(float(re.search(r"avg_precis?i?on:([0-9.e+-]+)%", line).group(1)))
            cur["rounded"] = "avg_precision:" in line
        elif line.startswith("Convergence_matrics:"):
            cur["convergence_line"] = float(line.split(":")[1])
        elif line.startswith("Convergence: ["):
            conv_lists.append([float(v) for v in line[len("Convergence: ["):-1].split(",")])
        elif line.startswith("Convergence reached"):
            converged = float(re.search(r"reached ([0-9.]+)%", line).group(1))
    # the convergence of epoch e is entry e of the list printed at its end (the last epoch prints none)
    last = conv_lists[-1] if conv_lists else []
    for e in epochs:
        i = e["epoch"]
        e["convergence"] = e.pop("convergence_line", None)
        if e["convergence"] is None and i < len(last) and last[i] != 0:
            e["convergence"] = last[i]
    return {"log": os.path.relpath(path, ROOT), "accuracy_threshold": thr, "ck": ck, "split": split,
            "epochs": epochs, "converged_percent": converged}


if __name__ == "__main__":
    ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    paths = sorted(glob.glob(os.path.join(ROOT, "data", "matrices_for_testing", "**",
                                          "*_search_resparse_*_CK*_split*.txt"), recursive=True))
    logs = [parse(p) for p in paths]
    logs = [l for l in logs if l["epochs"]]
    with open(os.path.join(HERE, "resparse_search_logs.json"), "w") as f:
        json.dump(logs, f, indent=1)
    print(f"{len(logs)} logs, {sum(len(l['epochs']) for l in logs)} epochs")
