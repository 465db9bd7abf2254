"""Writes solvertest_5x5.json: the only known-answer system the reference
holds (solvertest.f90:10-13): CSR row pointers (1-based) (1,4,6,8,11,14),
columns (1,3,4,2,4,1,3,2,3,4,1,2,5), all values 1, right-hand side
(8,6,4,9,8).  Its unique solution (1,2,3,4,5) follows by elimination:
x4 = 8-4 from rows 2 and 4 ... (x1 = 4-3, x2 = 6-4, x3 = 9-2-4, x5 = 8-1-2).
The reference's own driver cannot check it (its bicgstab multiplies by the
global fmatrix, solver.f90:163), so the data are kept as a fixture."""
import json
import os

rowptr1 = [1, 4, 6, 8, 11, 14]
col1 = [1, 3, 4, 2, 4, 1, 3, 2, 3, 4, 1, 2, 5]
data = {
    "source": "solvertest.f90:10-13",
    "rowptr": [r - 1 for r in rowptr1],
    "col": [c - 1 for c in col1],
    "val": [1.0] * 13,
    "b": [8.0, 6.0, 4.0, 9.0, 8.0],
    "x": [1.0, 2.0, 3.0, 4.0, 5.0],
    "guess": [1.0] * 5,
}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "solvertest_5x5.json"), "w") as f:
    json.dump(data, f, indent=1)
