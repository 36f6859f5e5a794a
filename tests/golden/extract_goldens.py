"""Writes tests/golden/reference_goldens.json from the golden vectors held in the
reference's own unit tests.  Run in the container that has the reference mounted:

    python tests/golden/extract_goldens.py /root/reference

Only numbers are extracted (test inputs are restated in tests/test_oracle_golden.py):
  unit_tests/DerivativeTests.cpp   3D (:51-264), ElastostaticResidual3D (:273-489),
                                   InternalElasticEnergy3D (:499-698), StressPNorm3D (:972-1140),
                                   Volume3D (:2151-2302)
  unit_tests/OmegaHUtilitiesTests.cpp  LocalElementCoords_3D (1x1x1 box element coordinates)
  unit_tests/PlatoAugLagStressTest.cpp AugLag_VonMises3D (:1063-1101): input stresses and
                                   von Mises values
"""
from __future__ import annotations

import json
import re
import sys
from pathlib import Path

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def numbers(txt: str) -> list[float]:
    return [float(x) for x in re.findall(NUM, txt)]


def test_body(src: str, name: str) -> str:
    m = re.search(r"TEUCHOS_UNIT_TEST\(\s*DerivativeTests,\s*" + re.escape(name) + r"\s*\)(.*?)(?=TEUCHOS_UNIT_TEST\()",
                  src, re.S)
    if not m:
        raise KeyError(name)
    return m.group(1)


def array(body: str, var: str) -> list[float]:
    m = re.search(re.escape(var) + r"\s*=\s*\{(.*?)\};", body, re.S)
    if not m:
        raise KeyError(var)
    return numbers(m.group(1))


def scalar(body: str, var: str) -> float:
    m = re.search(r"Plato::Scalar\s+" + re.escape(var) + r"\s*=\s*(" + NUM + r")\s*;", body)
    if not m:
        raise KeyError(var)
    return float(m.group(1))


def main(ref: Path) -> None:
    src = (ref / "unit_tests" / "DerivativeTests.cpp").read_text()
    out: dict = {"source": "unit_tests/DerivativeTests.cpp, unit_tests/OmegaHUtilitiesTests.cpp, "
                        "unit_tests/PlatoAugLagStressTest.cpp"}
    b = test_body(src, "3D")
    out["cells3d"] = {k: array(b, k) for k in ("gradient_gold", "strain_gold", "stress_gold", "result_gold")}
    b = test_body(src, "ElastostaticResidual3D")
    out["elastostatic_residual3d"] = {k: array(b, k)
                                      for k in ("residual_gold", "gold_jac_entries", "gold_grad_entries",
                                                "gold_grad_x_entries")}
    b = test_body(src, "InternalElasticEnergy3D")
    out["internal_elastic_energy3d"] = {"value_gold": scalar(b, "value_gold"),
                                        "grad_u_gold": array(b, "grad_u_gold"),
                                        "grad_z_gold": array(b, "grad_z_gold"),
                                        "grad_x_gold": array(b, "grad_x_gold")}
    b = test_body(src, "StressPNorm3D")
    out["stress_pnorm3d"] = {"value_gold": scalar(b, "value_gold"), "grad_u_gold": array(b, "grad_u_gold"),
                             "grad_z_gold": array(b, "grad_z_gold"), "grad_x_gold": array(b, "grad_x_gold")}
    b = test_body(src, "Volume3D")
    out["volume3d"] = {"value_gold": scalar(b, "value_gold"), "grad_z_gold": array(b, "grad_z_gold"),
                       "grad_x_gold": array(b, "grad_x_gold")}
    asrc = (ref / "unit_tests" / "PlatoAugLagStressTest.cpp").read_text()
    m = re.search(r"TEUCHOS_UNIT_TEST\(\s*PlatoAnalyzeUnitTests,\s*AugLag_VonMises3D\s*\)(.*?)(?=TEUCHOS_UNIT_TEST\()",
                  asrc, re.S)
    body = m.group(1)
    stress = {}
    for c, k, v in re.findall(r"tHostCauchyStress\((\d),\s*(\d)\)\s*=\s*(" + NUM + r")\s*;", body):
        stress[(int(c), int(k))] = float(v)
    ncell = 1 + max(c for c, _ in stress)
    gm = re.search(r"tGold\s*=\s*\{(.*?)\};", body, re.S)
    out["von_mises3d"] = {"stress": [[stress[(c, k)] for k in range(6)] for c in range(ncell)],
                          "gold": numbers(gm.group(1)), "tolerance": 1e-4}
    usrc = (ref / "unit_tests" / "OmegaHUtilitiesTests.cpp").read_text()
    m = re.search(r"LocalElementCoords_3D\)(.*?)TEUCHOS_UNIT_TEST", usrc, re.S)
    coords = {}
    for c, n, d, v in re.findall(r"tHostGold\((\d),(\d),(\d)\)\s*=\s*(" + NUM + r")", m.group(1)):
        coords[(int(c), int(n), int(d))] = float(v)
    out["local_element_coords_1x1x1"] = [[[coords[(c, n, d)] for d in range(3)] for n in range(4)]
                                         for c in range(6)]
    dst = Path(__file__).resolve().parent / "reference_goldens.json"
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
