"""Write treion_fit_coefficients.json: the post-front temperature fit of
set_Treion_default (gas_funcs.cc:104-129), read as text from the reference's
source (the Cn table at gas_funcs.cc:109; D'Aloisio et al. 2019).

Run in a container that has the reference checkout:
    python tests/golden/make_treion_fixture.py /root/reference
The fixture holds the five numbers only; tests read the JSON.
"""
import json
import os
import re
import sys


def extract(ref_dir):
    path = os.path.join(ref_dir, "gas_funcs.cc")
    with open(path) as f:
        lines = f.read().splitlines()
    start = next(i for i, l in enumerate(lines) if "void set_Treion_default" in l)
    for i in range(start, start + 20):
        m = re.search(r"float\s+Cn\s*\[\s*5\s*\]\s*=\s*\{([^}]*)\}", lines[i])
        if m:
            vals = [float(v) for v in m.group(1).split(",")]
            assert len(vals) == 5
            return {"source": "gas_funcs.cc:%d" % (i + 1), "Cn": vals,
                    "meaning": "ln T_reion[K] = sum_t Cn[t] * ln(v_IF / 1e5 cm/s)^t"}
    raise RuntimeError("Cn table not found after set_Treion_default")


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "treion_fit_coefficients.json")
    with open(out, "w") as f:
        json.dump(extract(ref), f, indent=1)
        f.write("\n")
    print("wrote", out)
