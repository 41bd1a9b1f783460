"""Writes the golden fixtures in this folder (run from anywhere; no reference
code is imported or executed -- the values below are transcribed from the
reference's test and README, and the 50x50 matrix is regenerated with the
standard-library procedure the reference's CI script describes)."""
import json
import os
import random

HERE = os.path.dirname(os.path.abspath(__file__))


def unit_grammar():
    # reference gpu-engine/cpu_selftest.cpp:378-414 and asserts :425-463
    return {
        "source": "gpu-engine/cpu_selftest.cpp:374-467",
        "alpha": 5,
        "kind": [0, 0, 1, 0],
        "left": [1, 3, 6, 7],
        "right": [2, 5, -1, 5],
        "rep": [1, 1, 3, 1],
        "C": [8, 0],
        "rows": 2,
        "cols": 3,
        "V": [1.5, 2.5],
        "x": [0.5, -0.5, 1.0],
        "expect_max_depth": 4,
        "expect_num_passthrough": 5,
        "expect_y": [4.5, 0.0],
    }


def readme_8x7():
    # reference README.md:151-183
    M = [
        [2.3, 0, 1.2, 1.2, 1.2, 2.3, 1.2],
        [1.2, 4.5, 3.4, 1.2, 1.2, 2.3, 4.5],
        [1.2, 4.5, 3.4, 1.2, 1.2, 2.3, 4.5],
        [2.3, 0, 1.2, 1.2, 1.2, 2.3, 1.2],
        [1.2, 4.5, 3.4, 1.2, 1.2, 2.3, 4.5],
        [0, 4.5, 3.4, 2.3, 2.3, 0, 4.5],
        [0, 4.5, 3.4, 2.3, 2.3, 0, 4.5],
        [2.3, 0, 1.2, 1.2, 1.2, 2.3, 1.2],
    ]
    return {"source": "README.md:145-183", "M": M, "x": [2, 1, 1, 2, 1, 1, 1],
            "expect_y": [12.9, 20.7, 20.7, 12.9, 20.7, 19.3, 19.3, 12.9]}


def integration_50x50():
    # reference tests/run_integration_tests.py:113-124: five patterns of 50
    # values drawn from {0,1,2}, then 50 rows each a random pattern.
    rows = cols = 50
    random.seed(42)
    patterns = [[float(random.choice([0, 1, 2])) for _ in range(cols)] for _ in range(5)]
    M = [list(random.choice(patterns)) for _ in range(rows)]
    return {"source": "tests/run_integration_tests.py:110-128", "rows": rows, "cols": cols, "M": M}


if __name__ == "__main__":
    for name, fn in [("unit_grammar", unit_grammar), ("readme_8x7", readme_8x7),
                     ("integration_50x50", integration_50x50)]:
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(fn(), f, indent=1)
        print("wrote", name)
