"""Write the reference's golden progress rows as JSON fixtures.

Reads tests/tests.ini of the reference checkout (three lines per case: name,
the singleideal.ini patch, the expected last row of progress.out) and keeps
the cases whose path this library covers.  Run once in a checkout that has
the reference: python tests/golden/extract_tests_ini.py /path/to/reference
"""
import json
import os
import sys

KEEP = {"subsonic_channel"}
COLUMNS = ["icyc", "dt", "time", "residual_rhou", "-dpdx/dt", "rhobulk", "ubulk", "tbulk",
           "rhomin", "rhomax", "tmin", "tmax", "pmin", "pmax"]


def main(ref):
    lines = [ln.strip() for ln in open(os.path.join(ref, "tests", "tests.ini")) if ln.strip()]
    here = os.path.dirname(os.path.abspath(__file__))
    i = 0
    while i < len(lines):
        if lines[i].startswith("["):
            name = lines[i][1:-1].strip()
            if name in KEEP:
                row = lines[i + 2].split()
                out = {"case": name, "source": "tests/tests.ini", "patch": lines[i + 1],
                       "columns": COLUMNS, "row_text": row, "row": [float(v) for v in row]}
                with open(os.path.join(here, f"{name}_progress.json"), "w") as f:
                    json.dump(out, f, indent=1)
            i += 3
        else:
            i += 1


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
