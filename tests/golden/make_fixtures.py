"""Convert the reference's golden data files into compact fixtures under tests/golden/.

The reference (amnp95/FemoraX) holds OpenSees output as %.17g text in
tests/reference/data/ (format: tests/reference/drivers/driver_common.hpp:1-9). This script
reads the files the hot path needs and stores them unchanged in value (float64 round-trips
%.17g exactly) as compressed .npz, so the tests never need the reference tree:

  brick.npz        element/Brick_*_{state,mass}.txt   (stdBrick + ElasticIsotropic, 16 files)
  ssp_brick.npz    element/SSPbrick_*_{state,mass}.txt (SSPbrick + ElasticIsotropic, 16 files)
  elastic.npz      nd/ElasticIsotropic_*.txt          (20 strain paths)
  pimy.npz         nd/PressureIndependMultiYield_*.txt (34 strain paths)
  constraints.npz  constraints/cases/*.txt + constraints/*.txt (OpenSees equation classes)
  system_pimy.npz  system/<case>/opensees_{disp,vel,accel}.txt for the PressureIndependMultiYield
                   soil-column cases (implicit stdBrick; SSPbrick with explicit shaking): OpenSees'
                   node recorders, one row per shaking step: time, then nodes x DOFs
  recorder_text.json  the recorders' text itself: the time tokens of
                   system/column25_ssp_explicit/opensees_disp.txt (-dT 0.01 at dt 2.5e-4) and the
                   first three lines of system/column4_pimy_std_implicit/opensees_disp.txt

Each data file becomes three arrays: "<name>/data" (rows x columns), "<name>/columns" and
"<name>/header" (the '# key value...' lines as a JSON object). Run:
    python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_reference_file(path):
    header, columns, rows = {}, [], []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\n")
            if not line:
                continue
            if line.startswith("#"):
                parts = line[1:].strip().split()
                if not parts:
                    continue
                if parts[0] == "columns":
                    columns = parts[1:]
                else:
                    header[parts[0]] = parts[1:]
                continue
            rows.append([float(t) for t in line.split()])
    data = np.array(rows, dtype=np.float64).reshape(len(rows), len(columns) if columns else -1)
    return header, columns, data


def pack(files, out):
    arrays = {}
    for path in files:
        name = os.path.basename(path)[:-4]
        header, columns, data = read_reference_file(path)
        arrays[name + "/data"] = data
        arrays[name + "/columns"] = np.array(columns)
        arrays[name + "/header"] = np.array(json.dumps(header))
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print(out, len(files), "files")


def read_case(path):
    nodes, fixes, ties, expect = [], [], [], ""
    commands = 0
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "node":
                nodes.append([int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5])])
            elif t[0] == "fix":
                fixes.append([int(t[1]), int(t[2])])
            elif t[0] == "equal":
                # one equalDOF command per "equal" line (the id is the line's ordinal)
                for d in t[3:]:
                    ties.append([int(t[1]), int(t[2]), int(d), commands])
                commands += 1
            elif t[0] == "expect":
                expect = " ".join(t[1:])
    return nodes, fixes, ties, expect


def pack_constraints(ref, out):
    d = os.path.join(ref, "tests/reference/data/constraints")
    arrays = {}
    names = sorted(n[:-4] for n in os.listdir(os.path.join(d, "cases")) if n.endswith(".txt"))
    for name in names:
        nodes, fixes, ties, expect = read_case(os.path.join(d, "cases", name + ".txt"))
        arrays[name + "/nodes"] = np.array(nodes, dtype=np.float64).reshape(-1, 5)
        arrays[name + "/fix"] = np.array(fixes, dtype=np.int32).reshape(-1, 2)
        arrays[name + "/equal"] = np.array(ties, dtype=np.int32).reshape(-1, 4)
        arrays[name + "/expect"] = np.array(expect)
        eqf = os.path.join(d, name + ".txt")
        if os.path.exists(eqf):
            _, cols, data = read_reference_file(eqf)
            arrays[name + "/opensees"] = data.astype(np.int64)
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print(out, len(names), "cases")


SYSTEM_CASES = ("element_std_clay_implicit", "column4_pimy_std_implicit", "element_ssp_frictional_explicit",
                "column4_pimy_ssp_explicit_newton", "column4_pimy_ssp_explicit_vtkhdf")


def pack_system(ref, out):
    arrays = {}
    for case in SYSTEM_CASES:
        for resp in ("disp", "vel", "accel"):
            path = os.path.join(ref, "tests/reference/data/system", case, f"opensees_{resp}.txt")
            arrays[f"{case}/{resp}"] = np.loadtxt(path, ndmin=2)
    np.savez_compressed(os.path.join(HERE, out), **arrays)
    print(out, len(SYSTEM_CASES), "cases")


def pack_recorder_text(ref, out):
    sysdir = os.path.join(ref, "tests/reference/data/system")
    with open(os.path.join(sysdir, "column25_ssp_explicit", "opensees_disp.txt")) as f:
        times = [line.split(" ", 1)[0] for line in f]
    with open(os.path.join(sysdir, "column4_pimy_std_implicit", "opensees_disp.txt")) as f:
        head = [f.readline() for _ in range(3)]
    doc = {"column25_ssp_explicit_times": {"dt": 2.5e-4, "dT": 0.01, "precision": 17, "times": times},
           "column4_pimy_std_implicit_disp_head": {"precision": 17, "lines": head}}
    with open(os.path.join(HERE, out), "w") as f:
        json.dump(doc, f, indent=1)
    print(out, len(times), "times", len(head), "lines")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = os.path.join(ref, "tests/reference/data")
    el = sorted(os.path.join(data, "element", f) for f in os.listdir(os.path.join(data, "element"))
                if f.startswith("Brick_"))
    pack(el, "brick.npz")
    ssp = sorted(os.path.join(data, "element", f) for f in os.listdir(os.path.join(data, "element"))
                 if f.startswith("SSPbrick_"))
    pack(ssp, "ssp_brick.npz")
    nd = os.path.join(data, "nd")
    pack(sorted(os.path.join(nd, f) for f in os.listdir(nd) if f.startswith("ElasticIsotropic_")), "elastic.npz")
    pack(sorted(os.path.join(nd, f) for f in os.listdir(nd) if f.startswith("PressureIndependMultiYield_")),
         "pimy.npz")
    pack_constraints(ref, "constraints.npz")
    pack_system(ref, "system_pimy.npz")
    pack_recorder_text(ref, "recorder_text.json")


if __name__ == "__main__":
    main()
