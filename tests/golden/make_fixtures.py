"""Writes tests/golden/reference_fixtures.json from the reference's own test
data (run once in the build container, where the reference tree is mounted;
the JSON is committed and nothing at test time reads the reference).

Extracted data (numbers only):
  * the five 8-spin instances and their stored ground states of
    tests/cxxtest/testcase.hpp (J(i,i) is the field h_i there);
  * the sampler fixtures of tests/test_sampler.py:10-47 (h/J model with ground
    energy -6, QUBO with ground energy -3, 30-spin 1D antiferromagnet);
  * the layout matrix of tests/test_model.py:57-68;
  * the Trotter get_solution case of tests/cxxtest/result/result.hpp:21-41;
  * the EnergyCheck size of tests/cxxtest/graph/quadratic.hpp:129-195.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")


def parse_testcases(text):
    cases = {}
    blocks = re.split(r"#(?:el)?if TEST_CASE_INDEX == (\d+)", text)
    for k in range(1, len(blocks), 2):
        idx = int(blocks[k])
        body = blocks[k + 1]
        entries = re.findall(r"interaction\.J\((\d+),\s*(\d+)\)\s*=\s*([-+0-9.eE]+);", body)
        gs = re.search(r"Spins\(\{([-0-9,\s]+)\}\)", body)
        hard = "Hard" in body
        n = 8
        J = [[0.0] * (n + 1) for _ in range(n + 1)]
        for i, j, v in entries:
            i, j, v = int(i), int(j), float(v)
            if i == j:
                J[i][n] = v  # J(i,i) is h_i
            else:
                J[min(i, j)][max(i, j)] = v
        J[n][n] = 1.0
        cases[str(idx)] = {
            "n": n,
            "J_ext_upper": J,
            "ground_state": [int(x) for x in gs.group(1).split(",")],
            "hard": hard,
        }
    return cases


def main():
    with open(os.path.join(REF, "tests/cxxtest/testcase.hpp")) as f:
        tc = f.read()
    m = re.search(r"make_classical_schedule_list\(([^)]*)\)", tc)
    cl = [float(x) for x in m.group(1).split(",")]
    m = re.search(r"make_transverse_field_schedule_list\(([^)]*)\)", tc)
    tr = [float(x) for x in m.group(1).split(",")]
    fixtures = {
        "source": "tests/cxxtest/testcase.hpp, tests/test_sampler.py, tests/test_model.py, "
                  "tests/cxxtest/result/result.hpp, tests/cxxtest/graph/quadratic.hpp",
        "testcases": parse_testcases(tc),
        "classical_schedule": {"beta_min": cl[0], "beta_max": cl[1], "one_mc_step": int(cl[2]),
                               "num_call_updater": int(cl[3])},
        "transverse_schedule": {"beta": tr[0], "one_mc_step": int(tr[1]),
                                "num_call_updater": int(tr[2]), "gamma": 1.0,
                                "num_trotter": 10},
        "sampler_ising": {"h": {"0": -1, "1": -1, "2": 1, "3": 1},
                          "J": [[0, 1, -1], [3, 4, -1]],
                          "ground_state": [1, 1, -1, -1, -1], "ground_energy": -6.0},
        "sampler_qubo": {"Q": [[0, 0, -1], [1, 1, -1], [2, 2, 1], [3, 3, 1], [4, 4, 1],
                               [0, 1, -1], [3, 4, 1]],
                         "ground_state": [1, 1, 0, 0, 0], "ground_energy": -3.0},
        "sampler_afm": {"N": 30, "h": {"0": -10}, "J": 1.0,
                        "ground_state": [(-1) ** i for i in range(30)]},
        "model_layout": {"h": {"0": 1, "1": -2}, "J": [[0, 1, -1], [1, 2, -3], [2, 3, 0.5]],
                         "matrix": [[0, -1, 0, 0, 1], [0, 0, -3, 0, -2], [0, 0, 0, 0.5, 0],
                                    [0, 0, 0, 0, 0], [0, 0, 0, 0, 1]]},
        "trotter_solution": {"n": 4, "h": {"1": -1.0},
                             "J": [[0, 1, -1.0], [1, 2, -1.0], [2, 3, 1.0]],
                             "slice0": [1, 1, 1, -1], "num_trotter": 4},
        "energy_check": {"N": 500},
    }
    with open(OUT, "w") as f:
        json.dump(fixtures, f, indent=1)
    print("wrote", OUT, "cases:", sorted(fixtures["testcases"]))


if __name__ == "__main__":
    main()
