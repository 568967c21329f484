#!/usr/bin/env python3
"""Write tests/golden/example_params.json: the key = value pairs of the reference's example
cases (tutorials/examples/Example0*.in: neutral, convective and stable boundary layers and
BOMEX), comments dropped, in file order.  The parameter values are data; the test
tests/test_example_params.py feeds them through this build's .in parser and host set-up.

Run from the repository root where the reference tree is mounted:
    python3 tests/golden/make_example_params.py
"""
import glob
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/tutorials/examples"


def pairs(path):
    out = []
    for line in open(path):
        body = line.split("#", 1)[0]
        if "=" not in body:
            continue
        k, v = body.split("=", 1)
        out.append([k.strip(), v.strip()])
    return out


def main():
    cases = {os.path.basename(f): pairs(f) for f in sorted(glob.glob(os.path.join(SRC, "Example0*.in")))}
    with open(os.path.join(HERE, "example_params.json"), "w") as fh:
        json.dump(cases, fh, indent=1)
    print({k: len(v) for k, v in cases.items()})


if __name__ == "__main__":
    main()
