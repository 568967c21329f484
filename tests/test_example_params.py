"""The reference's example cases (tutorials/examples/Example0*.in, parameter values committed as
tests/golden/example_params.json) through this build's .in parser and host set-up.

Each case keeps every parameter of the example except the ones that would need the reference's
4-rank decomposition or its netCDF input: one rank, a 32 x 32 horizontal grid (the vertical
grid, stretching and physics are the example's), and no inFile for BOMEX (whose initial file is
a netCDF download).  The set-up must succeed without parameter errors, every selector and
constant it reads must hold the example's value, and the parameters it leaves unread must be
ones the reference also leaves unread under the example's selectors (moisture keys without
moisture, large-scale forcing keys without forcing, the idealised-sine surface keys without
that forcing, the free-text Description).
"""
import json
import os

import numpy as np
import pytest

import fe_py

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = json.load(open(os.path.join(HERE, "golden", "example_params.json")))
OVERRIDE = {"numProcsX": "1", "numProcsY": "1", "Nx": "32", "Ny": "32", "inFile": "", "inPath": ""}


def case_params(name, tmp_path):
    out = []
    for k, v in CASES[name]:
        if k in ("outPath",):
            v = str(tmp_path) + "/"
        out.append((k, OVERRIDE.get(k, v)))
    return out


def write_case(pairs, path):
    with open(path, "w") as fh:
        for k, v in pairs:
            fh.write(f"{k} = {v}\n")


def allowed_unread(values):
    sel = {k: v for k, v in values}
    ok = {"Description"}
    if int(sel.get("moistureSelector", "0")) == 0:
        ok |= {k for k, _ in values if k.startswith("moisture") and k != "moistureSelector"}
        ok |= {"surflayer_wq", "surflayer_qr", "surflayer_ideal_qts", "surflayer_ideal_qte", "surflayer_ideal_qamp"}
    if int(sel.get("lsfSelector", "0")) == 0:
        ok |= {k for k, _ in values if k.startswith("lsf_")}
    if int(sel.get("surflayer_idealsine", "0")) == 0:
        ok |= {k for k, _ in values if k.startswith("surflayer_ideal_")}
    return ok


@pytest.mark.parametrize("name", sorted(CASES))
def test_reference_example_sets_up(name, tmp_path, capfd):
    pairs = case_params(name, tmp_path)
    path = os.path.join(str(tmp_path), "case.in")
    write_case(pairs, path)
    L = fe_py.Model.lib()
    capfd.readouterr()
    rc = L.fe_setup_host(path.encode())
    try:
        assert rc == 0 and L.getParameterErrors() == 0, (name, rc)
        libc = fe_py.C.CDLL(None)
        libc.fflush(None)
        capfd.readouterr()
        L.printUnusedParameters()
        libc.fflush(None)                     # the listing goes through C stdio
        unread = {line.split("=", 1)[0].strip() for line in capfd.readouterr().out.splitlines() if "NOT USED" in line}
        assert unread <= allowed_unread(pairs), sorted(unread - allowed_unread(pairs))
        checked = 0
        for k, v in pairs:
            if k in OVERRIDE or k in unread or v == "":
                continue
            iv, fv = fe_py.C.c_int(), fe_py.C.c_float()
            if L.fe_get_int(k.encode(), fe_py.C.byref(iv)) == 0:
                assert iv.value == int(v), (k, iv.value, v)
                checked += 1
            elif L.fe_get_float(k.encode(), fe_py.C.byref(fv)) == 0:
                assert fv.value == np.float32(float(v)), (k, fv.value, v)
                checked += 1
        print(f"{name}: {checked} parameters read back, {len(unread)} unread")
        assert checked >= 50, checked
    finally:
        L.fe_teardown()
