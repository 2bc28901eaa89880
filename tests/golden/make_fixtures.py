"""Writes tests/golden/h2o_ccsd.json from the reference's committed validation
outputs (run in the development container, where the reference tree is
mounted; the tests only read the JSON).

Data taken: the h2o.xyz geometry, RHF energies, CCSD iteration tables
(iter, deltaE, residual, energy) and final CCSD energies of
  tests/validation/reference/outputs/h2o-eom-ccsd-direct-631g.out  (6-31G, frozen core)
  tests/validation/reference/outputs/h2o-eom-ip-ccsd-pvdz.out      (cc-pVDZ, all electron)
  tests/validation/reference/outputs/h2o-eom-ea-ccsd-pvdz.out      (cc-pVDZ, frozen core)
and, under "df_cases", the density-fitted run
  tests/validation/reference/outputs/h2o-ccsd-631g-pvdz.out        (RI-RHF 6-31G with the
      cc-pVDZ fitting basis, CCSD method "df", reduced_abcd_memory, frozen core,
      precision 1e-11)
under "eom_cases", the converged EOM-CCSD singlet excitation energies (last Davidson
table, 12 digits) of
  tests/validation/reference/outputs/h2o-eom-ccsd-direct-631g.out  (6-31G, frozen core)
  tests/validation/reference/outputs/h2o-eom-ccsd-631g-pvdz.out    (DF: 6-31G / cc-pVDZ fit)
and, under "triples_cases", the (T) energy of
  tests/validation/reference/outputs/h2o-ccsd_t-631g-pvdz.out      (the same DF water,
      CCSD(T) with approach "straight")
"""
import json
import os
import re

REF = "/root/reference/tests/validation/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse(name):
    txt = open(os.path.join(REF, "outputs", name)).read()
    rev = re.search(r"Revision (\w+)", txt).group(1)
    scf = [float(x) for x in re.findall(r"Energy: (-?\d+\.\d+)", txt)]
    rows = []
    m = re.search(r"iter \t +deltaE.*?\n(.*?)CCSD Energy", txt, re.S)
    for line in m.group(1).strip().splitlines():
        p = line.split()
        rows.append({"iter": int(p[0]), "dE": float(p[1]), "residual": float(p[2]),
                     "energy": float(p[3])})
    e = float(re.search(r"CCSD Energy\s+(-?\d+\.\d+)", txt).group(1))
    return rev, scf[-1], rows, e


cases = [
    ("h2o-eom-ccsd-direct-631g.out", "6-31G", True, 4, 8, 1e-9, 1.0,
     "current convergence rule: dE <= 0.1*precision and error <= precision"),
    ("h2o-eom-ip-ccsd-pvdz.out", "cc-pVDZ", False, 5, 19, 1e-9, 10.0,
     "revision 7efc976 stops at dE <= precision (no 0.1 margin): equivalent to the current "
     "rule with precision x10 for this trace"),
    ("h2o-eom-ea-ccsd-pvdz.out", "cc-pVDZ", True, 4, 19, 1e-9, 10.0,
     "revision 7efc976 stops at dE <= precision (no 0.1 margin): equivalent to the current "
     "rule with precision x10 for this trace"),
]
out = {"geometry_angstrom": open(os.path.join(REF, "inputs", "h2o.xyz")).read(),
       "units": "2010CODATA", "cases": []}
for fname, basis, fc, o, v, prec, scale, note in cases:
    rev, e_scf, rows, e = parse(fname)
    out["cases"].append({"source": f"tests/validation/reference/outputs/{fname}", "revision": rev,
                         "basis": basis, "frozen_core": fc, "o": o, "v": v, "e_rhf": e_scf,
                         "e_ccsd": e, "precision": prec, "precision_scale_for_current_rule": scale,
                         "note": note, "trace": rows})
out["df_cases"] = []
for fname, basis, dfbs, fc, o, v, prec in [("h2o-ccsd-631g-pvdz.out", "6-31G", "cc-pVDZ", True, 4, 8,
                                            1e-11)]:
    rev, e_scf, rows, e = parse(fname)
    out["df_cases"].append({"source": f"tests/validation/reference/outputs/{fname}", "revision": rev,
                            "basis": basis, "df_basis": dfbs, "method": "df",
                            "reduced_abcd_memory": True, "frozen_core": fc, "o": o, "v": v,
                            "e_rhf": e_scf, "e_ccsd": e, "precision": prec,
                            "precision_scale_for_current_rule": 1.0,
                            "note": "RI-RHF reference; current convergence rule", "trace": rows})
out["triples_cases"] = []
for fname in ["h2o-ccsd_t-631g-pvdz.out"]:
    txt = open(os.path.join(REF, "outputs", fname)).read()
    e_t = float(re.search(r"\(T\) Energy: (-?\d+\.\d+(?:e-?\d+)?)", txt).group(1))
    e_cc = float(re.search(r"CCSD Energy\s+(-?\d+\.\d+)", txt).group(1))
    out["triples_cases"].append({"source": f"tests/validation/reference/outputs/{fname}",
                                 "revision": re.search(r"Revision (\w+)", txt).group(1),
                                 "basis": "6-31G", "df_basis": "cc-pVDZ", "method": "df",
                                 "frozen_core": True, "approach": "straight", "e_ccsd": e_cc,
                                 "e_triples": e_t})
out["eom_cases"] = []
for fname, basis, dfbs in [("h2o-eom-ccsd-direct-631g.out", "6-31G", None),
                           ("h2o-eom-ccsd-631g-pvdz.out", "6-31G", "cc-pVDZ")]:
    txt = open(os.path.join(REF, "outputs", fname)).read()
    head = txt.split("Excitation Energy: ( Singlets )")[0]
    last = head.split("iter=")[-1]
    roots = [float(m.group(1)) for m in re.finditer(r"^\s+\d+\s+\S+\s+\S+\s+(\d+\.\d+)\s*$", last, re.M)]
    table = txt.split("Excitation Energy: ( Singlets )")[1]
    summary = [float(m.group(1)) for m in re.finditer(r"^\s+\d+\s+(\d+\.\d+)\s+\d+\.\d+\s+\d+\.\d+",
                                                       table.split("EOM-CCSD Total Time")[0], re.M)]
    out["eom_cases"].append({"source": f"tests/validation/reference/outputs/{fname}",
                             "revision": re.search(r"Revision (\w+)", txt).group(1), "basis": basis,
                             "df_basis": dfbs, "frozen_core": True, "singlets": roots,
                             "singlets_summary": summary})
with open(os.path.join(HERE, "h2o_ccsd.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", len(out["cases"]), "cases,", len(out["df_cases"]), "DF cases,", len(out["triples_cases"]),
      "(T) cases,", len(out["eom_cases"]), "EOM cases")
