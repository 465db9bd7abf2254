"""Writes fortran_records.json: records of six doubles and the lines a
Fortran run-time library prints for them with the edit descriptors of the
reference's text output (output.f90:205-222 snapshot records,
output.f90:162 probe series).

Compiles fortran_records.f90 (in this folder) with the image's flang and
runs it on the inputs; the JSON keeps the inputs as hex doubles and the
printed lines.  Run from the repository root:

    python tests/golden/make_fortran_records.py
"""
import json
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def inputs():
    rng = np.random.default_rng(20261202)
    special = [0.0, -0.0, 0.0625, -0.0625, 0.1875, 0.0004, -0.0004, 0.0005, 0.0015, 0.0025,
               9.9995, 9.99951, 1.0625, 2.5e-4, 999999.9994, 999999.9996, -99999.9994,
               -999999.9994, 1234567.0, 1e-310, -1e-310, 1e300, -1e300, 9.9995e99, 1e100,
               9.9994999e-100, float("inf"), float("-inf"), float("nan"), 0.5, -0.5, 1.0,
               0.99999, 0.999995, 0.0000049999, 0.0000050001, 12345.678949, 12345.67895]
    recs = []
    sp = list(special)
    while sp:
        chunk = (sp[:6] + [0.0] * 6)[:6]
        recs.append(chunk)
        recs.append(chunk[::-1])
        sp = sp[6:]
    for _ in range(200):
        x = rng.uniform(-0.01, 2.0)
        y = rng.uniform(-0.01, 1.0)
        p = rng.uniform(-2e3, 2e4) * (10.0 ** rng.integers(-3, 3))
        vx = rng.choice([-1, 1]) * 10.0 ** rng.uniform(-12, 0.6)
        vy = rng.choice([-1, 1]) * 10.0 ** rng.uniform(-320, 300)
        c = rng.uniform(0.0, 1.0)
        recs.append([x, y, p, vx, vy, c])
    return recs


def main():
    fc = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    recs = inputs()
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "fortran_records")
        subprocess.run([fc, "-O1", os.path.join(HERE, "fortran_records.f90"), "-o", exe],
                       check=True, cwd=d)
        fin = os.path.join(d, "in.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<i", len(recs)))
            for r in recs:
                f.write(struct.pack("<6d", *r))
        out = subprocess.run([exe, fin], check=True, capture_output=True, text=True).stdout
    lines = out.split("\n")[:-1]
    assert len(lines) == 2 * len(recs)
    data = {"compiler": subprocess.run([fc, "--version"], capture_output=True,
                                       text=True).stdout.splitlines()[0],
            "records": [{"v": [struct.pack("<d", x).hex() for x in r],
                         "snapshot": lines[2 * k], "probe": lines[2 * k + 1]}
                        for k, r in enumerate(recs)]}
    with open(os.path.join(HERE, "fortran_records.json"), "w") as f:
        json.dump(data, f, indent=0)


if __name__ == "__main__":
    main()
