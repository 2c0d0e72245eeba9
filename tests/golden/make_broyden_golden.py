"""Writes ref_broyden_h2o.json: the regression vectors of the reference's Broyden mixer unit test
(test/src/dftbp/unit/mixer/broyden.F90:25-64 -- a water molecule's charges and charge differences
over four SCC iterations, with the mixer parameters of that test and its tolerance).

Reads the numbers as data from the reference test file (only in the build container, where the
reference sources lie); the JSON it writes is what the tests load.
"""
import json
import os
import re

SRC = "/root/reference/test/src/dftbp/unit/mixer/broyden.F90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_broyden_h2o.json")


def fortran_array(text, name):
    m = re.search(name + r"\((\d+), *(\d+)\) *= *reshape\(\[(.*?)\]", text, re.S)
    rows, cols = int(m.group(1)), int(m.group(2))
    body = m.group(3).replace("&", " ").replace("_dp", "")
    vals = [float(v.replace("E", "e")) for v in re.findall(r"[-+]?\d+\.\d+(?:[Ee][-+]?\d+)?", body)]
    assert len(vals) == rows * cols, (name, len(vals))
    # Fortran column-major (rows, cols): column k is iteration k
    return [vals[k * rows:(k + 1) * rows] for k in range(cols)]


def main():
    text = open(SRC).read()
    params = {}
    for key in ("mixParam", "omega0", "minWeight", "maxWeight", "weightFac"):
        params[key] = float(re.search(r"inp%" + key + r" *= *([-0-9.Ee+]+)_dp", text).group(1))
    params["maxSccIter"] = int(re.search(r"inp%maxSccIter *= *(\d+)", text).group(1))
    data = {"source": "test/src/dftbp/unit/mixer/broyden.F90:25-64",
            "atol_eps": 100.0,
            "params": params,
            "expected_charges": fortran_array(text, "expectedCharges"),
            "difference": fortran_array(text, "difference")}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
