"""Writes the golden fixtures of tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_harness (this project's harness linked against the
reference's own hot-path sources) on each case of cases.py and stores, per
case, the reference's list_1 / list_2_1 / list_2_2 / list_Diagonal and one
H·v of the SplitMix64 vector with SEED, plus prdct:

* golden.json       dims, SHA-256 of every table, prdct, H·v checksums
* <case>.npz        the full arrays for cases with dim <= FULL_MAX

Full-size workloads (pass --large) store checksums only: S1 = sum v0,
S2 = sum |v0|^2, S3 = sum ((j mod 97)+1) v0[j] and 64 sampled entries
(the harness's `sums` mode).

Two-body terms (Hubbard InterAll, PairHop, Exchange; spin and SpinGC
InterAll; pass --two-body) are stored under "cases_two_body": the C
restatement in oracle/ does not cover them, so these fixtures come from the
reference alone.

Expectation values (pass --green) of the cases of green_cases() are stored
under "cases_green": energy, variance and moments (expec_energy_flct) and the
one- and two-body Green's functions (expec_cisajs, expec_cisajscktaltdc) of
the normalised synthetic vector.

Reference-written output files (pass --files) go to tests/golden/files/<case>/:
the Green's-function files of the green cases and the SS_rand / Norm_rand /
Flct_rand files of tpq_file_cases() (the reference's own CalcByTPQ) and of
ctpq_file_cases() (CalcByCanonicalTPQ), with each sample's start vector as the
reference made it (ini_rand<i>.bin: global_1st_norm, then the normalised v1);
the file writers are checked byte for byte against them and the TPQ
restatements against their numbers.

Block columns (pass --large-block): the same checksums for each of the 16
synthetic vectors of seeds 0..15 of the triangular 13x2 workload (harness mode
`sums_seeds`, one reference mltply per seed), i.e. the columns of the k = 16
block its mTPQ runs.

SpinGC three-, four- and six-body Green's functions (pass --nbody): the same
green mode on nbody_green_cases(); the ThreeBody / FourBody / SixBody lines
and the files the reference wrote (tests/golden/files/<case>/).

usage: python tests/golden/gen_golden.py [--large | --large-block | --two-body | --green | --nbody | --files]
"""

from __future__ import annotations

import hashlib
import json
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from hphi_amd import lattice as L  # noqa: E402
from oracle import oracle as O  # noqa: E402
from cases import SEED, ctpq_file_cases, golden_cases, green_cases, nbody_green_cases, tpq_file_cases, two_body_cases  # noqa: E402,E501
import shutil  # noqa: E402

FULL_MAX = 5000


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run_case(name, spec, work):
    d = os.path.join(work, name)
    L.write_expert(spec, d)
    O.run_reference(d, "tables", d + "/out")
    O.run_reference(d, "hv", d + "/out", seed=SEED)
    meta = dict(line.split() for line in open(d + "/out/meta.txt"))
    dim = int(meta["dim"])
    rec = {"dim": dim, "model": spec.model, "nsite": spec.nsite, "seed": SEED}
    arrays = {}
    canon = spec.model in (L.SPIN, L.HUBBARD)
    if canon:
        for t in ("list_1", "list_2_1", "list_2_2"):
            a = np.fromfile(f"{d}/out/{t}.bin", np.uint64)
            rec[f"sha_{t}"] = sha(a)
            rec[f"len_{t}"] = int(a.size)
            arrays[t] = a
    diag = np.fromfile(f"{d}/out/list_Diagonal.bin", np.float64)
    rec["sha_list_Diagonal"] = sha(diag)
    arrays["list_Diagonal"] = diag
    v0 = np.fromfile(f"{d}/out/v0.bin", np.complex128)
    pr = open(d + "/out/prdct.txt").readline().split()
    rec["prdct"] = [float(pr[0]), float(pr[1])]
    rec["hv_S1"] = [float(v0.sum().real), float(v0.sum().imag)]
    rec["hv_S2"] = float(np.sum(np.abs(v0) ** 2))
    rec["hv_max_abs"] = float(np.max(np.abs(v0)))
    arrays["hv"] = v0
    if dim <= FULL_MAX:
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        rec["arrays"] = f"{name}.npz"
    else:
        idx = (np.arange(64, dtype=np.uint64) * np.uint64(2654435761)) % np.uint64(dim)
        rec["hv_samples"] = [[int(j), float(v0[j].real), float(v0[j].imag)] for j in idx]
    return rec


def run_green(name, spec, work):
    """The reference's expectation values of the normalised SplitMix64 vector
    (seed SEED): expec_energy_flct's energy, variance and moments, and the
    OneBodyG / TwoBodyG lines of output/zvo_cisajs.dat, zvo_cisajscktalt.dat
    (values printed with 10 decimals)."""
    d = os.path.join(work, name)
    L.write_expert(spec, d)
    O.run_reference(d, "green", d + "/out", seed=SEED)
    phys = {k: float(v) for k, v in (ln.split() for ln in open(d + "/out/phys.txt") if ln.strip())}

    def lines(fname, nidx):
        path = os.path.join(d, "output", fname)
        if not os.path.exists(path):
            return []
        return [[int(x) for x in p[:nidx]] + [float(p[nidx]), float(p[nidx + 1])]
                for p in (ln.split() for ln in open(path)) if p]
    rec = {"dim": spec.dim, "seed": SEED, "phys": phys,
           "green1": lines("zvo_cisajs.dat", 4), "green2": lines("zvo_cisajscktalt.dat", 8)}
    for nb, fname in ((3, "zvo_ThreeBody.dat"), (4, "zvo_FourBody.dat"), (6, "zvo_SixBody.dat")):
        if getattr(spec, f"green{nb}"):
            # a line holds every index and the value, or eight indices and zeros
            path = os.path.join(d, "output", fname)
            rec[f"green{nb}"] = [[int(x) for x in p[:-2]] + [float(p[-2]), float(p[-1])]
                                 for p in (ln.split() for ln in open(path)) if p]
            keep_files(name, os.path.join(d, "output"), [fname])
    return rec


FILES = os.path.join(HERE, "files")


def keep_files(name, src_dir, names):
    """Copy reference-written output files into tests/golden/files/<name>/."""
    dst = os.path.join(FILES, name)
    os.makedirs(dst, exist_ok=True)
    kept = []
    for f in names:
        if os.path.exists(os.path.join(src_dir, f)):
            shutil.copyfile(os.path.join(src_dir, f), os.path.join(dst, f))
            kept.append(f)
    return kept


def run_files(work):
    """Reference-written output files: the Green's-function files of the
    green cases and the TPQ files of tpq_file_cases() (CalcByTPQ) and
    ctpq_file_cases() (CalcByCanonicalTPQ)."""
    rec = {}
    for name, spec in green_cases().items():
        d = os.path.join(work, "f_" + name)
        L.write_expert(spec, d)
        O.run_reference(d, "green", d + "/out", seed=SEED)
        rec[name] = {"model": spec.model,
                     "files": keep_files(name, d + "/output", ["zvo_cisajs.dat", "zvo_cisajscktalt.dat"])}
    for cases, ct, mode in ((tpq_file_cases(), L.CALC_TPQ, "tpq"), (ctpq_file_cases(), L.CALC_CTPQ, "ctpq")):
        for name, spec in cases.items():
            d = os.path.join(work, "f_" + name)
            L.write_expert(spec, d, calc_type=ct)
            O.run_reference(d, mode, d + "/out")
            nave = int(spec.params["NumAve"])
            names = [f"{k}_rand{i}.dat" for i in range(nave) for k in ("SS", "Norm", "Flct")]
            kept = keep_files(name, d + "/output", names)
            kept += keep_files(name, d + "/out", [f"ini_rand{i}.bin" for i in range(nave)])
            rec[name] = {"model": spec.model, "files": kept, "dim": spec.dim, "nsite": spec.nsite,
                         "large_value": float(spec.params["LargeValue"]),
                         "expand_coef": int(spec.params.get("ExpandCoef", 0)) or 10}
    return rec


def run_large(n, work):
    spec = L.baseline_config(n)
    d = os.path.join(work, f"cfg{n}")
    L.write_expert(spec, d)
    seed = {2: 1, 3: 2, 4: 0}[n]
    O.run_reference(d, "sums", d + "/out", seed=seed, timeout=6 * 3600)
    lines = open(d + "/out/sums.txt").read().split("\n")
    pr = open(d + "/out/prdct.txt").read().split()
    rec = {"dim": spec.dim, "seed": seed, "prdct": [float(pr[0]), float(pr[1])], "samples": []}
    for ln in lines:
        p = ln.split()
        if not p:
            continue
        if p[0] == "S1":
            rec["S1"] = [float(p[1]), float(p[2])]
        elif p[0] == "S2":
            rec["S2"] = float(p[1])
        elif p[0] == "S3":
            rec["S3"] = [float(p[1]), float(p[2])]
        elif p[0] == "E":
            rec["samples"].append([int(p[1]), float(p[2]), float(p[3])])  # 1-based index
    return rec


def run_large_block(n, k, work):
    """Reference checksums of H v_s for the k synthetic vectors of seeds
    s0 .. s0+k-1: the columns of a k-column block filled from seed s0."""
    spec = L.baseline_config(n)
    d = os.path.join(work, f"cfg{n}_k{k}")
    L.write_expert(spec, d)
    s0 = {2: 1, 3: 2, 4: 0}[n]
    O.run_reference(d, "sums_seeds", d + "/out", seed=s0, reps=k, timeout=6 * 3600)
    cols = []
    for s in range(s0, s0 + k):
        pr = open(f"{d}/out/prdct_{s}.txt").read().split()
        rec = {"seed": s, "prdct": [float(pr[0]), float(pr[1])], "samples": []}
        for ln in open(f"{d}/out/sums_{s}.txt"):
            p = ln.split()
            if not p:
                continue
            if p[0] == "S1":
                rec["S1"] = [float(p[1]), float(p[2])]
            elif p[0] == "S2":
                rec["S2"] = float(p[1])
            elif p[0] == "S3":
                rec["S3"] = [float(p[1]), float(p[2])]
            elif p[0] == "E":
                rec["samples"].append([int(p[1]), float(p[2]), float(p[3])])  # 1-based index
        cols.append(rec)
    return {"dim": spec.dim, "seed": s0, "k": k, "columns": cols}


def main():
    out_path = os.path.join(HERE, "golden.json")
    gold = json.load(open(out_path)) if os.path.exists(out_path) else {}
    with tempfile.TemporaryDirectory() as work:
        if "--two-body" in sys.argv:
            for name, spec in two_body_cases().items():
                gold.setdefault("cases_two_body", {})[name] = run_case(name, spec, work)
                print(name, "done", flush=True)
        elif "--files" in sys.argv:
            gold["files"] = run_files(work)
            print("files done", flush=True)
        elif "--nbody" in sys.argv:
            for name, spec in nbody_green_cases().items():
                gold.setdefault("cases_green", {})[name] = run_green(name, spec, work)
                print(name, "done", flush=True)
        elif "--green" in sys.argv:
            for name, spec in green_cases().items():
                gold.setdefault("cases_green", {})[name] = run_green(name, spec, work)
                print(name, "done", flush=True)
        elif "--large-block" in sys.argv:
            n, k = 4, 16
            gold.setdefault("large", {})[f"cfg{n}_k{k}"] = run_large_block(n, k, work)
            print("large block", n, k, "done", flush=True)
        elif "--large" in sys.argv:
            for n in [int(a) for a in sys.argv[sys.argv.index("--large") + 1:] if a.isdigit()] or [4, 2, 3]:
                gold.setdefault("large", {})[f"cfg{n}"] = run_large(n, work)
                json.dump(gold, open(out_path, "w"), indent=1)
                print("large", n, "done", flush=True)
        else:
            cases = {}
            for name, spec in golden_cases().items():
                cases[name] = run_case(name, spec, work)
                print(name, cases[name]["dim"], flush=True)
            gold["cases"] = cases
    json.dump(gold, open(out_path, "w"), indent=1)


if __name__ == "__main__":
    main()
