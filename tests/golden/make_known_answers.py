"""Write tests/golden/known_answers.json: the reference's own known-answer cases as data.

Inputs and expected `info` come from the reference tests:
  delayed_lu_cuda/example.cu:116-143  (dense 4x4, rank-3 6x6, rectangular 5x3, 70x70 far pivot)
  delayed_lu/example.cpp:204-253      (4096 x 4096, diag(2..801) = 800..1, info = 800)
Expected ipiv/jpiv follow from the algorithm by hand (each step's column is either the column
max or, for a zero column, the global max of the active block at a known position); the oracle
and the GPU path are both checked against them.  Run: python tests/golden/make_known_answers.py
"""
import json
import os

cases = []

dense = [[8.0, 1.0, 2.0, 0.0],
         [2.0, 9.0, 1.0, 1.0],
         [1.0, 1.0, 7.0, 2.0],
         [0.0, 1.0, 1.0, 6.0]]  # rows of the 4x4 whose columns are listed in example.cu:116-121
cases.append({"name": "dense square", "m": 4, "n": 4, "dense_rows": dense,
              "info": 4, "ipiv": [0, 1, 2, 3], "jpiv": [0, 1, 2, 3]})

cases.append({"name": "delayed rank-3", "m": 6, "n": 6,
              "entries": [[2, 2, 7.0], [4, 4, -5.0], [5, 5, 3.0]],
              "info": 3, "ipiv": [2, 4, 5, 3, 4, 5], "jpiv": [2, 4, 5, 3, 4, 5]})

rect = [[5.0, 1.0, 0.0],
        [1.0, 6.0, 1.0],
        [0.0, 1.0, 7.0],
        [1.0, 0.0, 1.0],
        [0.0, 1.0, 0.0]]
cases.append({"name": "rectangular 5x3", "m": 5, "n": 3, "dense_rows": rect,
              "info": 3, "ipiv": [0, 1, 2], "jpiv": [0, 1, 2]})

far_ipiv = list(range(70))
far_ipiv[1] = 69
cases.append({"name": "mid-panel far pivot", "m": 70, "n": 70,
              "entries": [[0, 0, 100.0], [69, 69, 5.0]],
              "info": 2, "ipiv": far_ipiv, "jpiv": far_ipiv})

big = 4096
ipiv = [k + 2 for k in range(800)] + list(range(800, big))
cases.append({"name": "example.cpp rank 800", "m": big, "n": big,
              "entries": [[k + 2, k + 2, float(800 - k)] for k in range(800)],
              "info": 800, "ipiv": ipiv, "jpiv": ipiv})

out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "known_answers.json")
with open(out, "w") as f:
    json.dump({"source": "reference test inputs (example.cu, example.cpp); expected pivots derived "
                         "by hand from the algorithm",
               "cases": cases}, f)
print(out)
