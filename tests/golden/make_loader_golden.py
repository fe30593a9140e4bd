"""Write tests/golden/loader_golden.json from the reference's own robot loader.

The approximate gate model (one bounding sphere per joint and the approximate
self-collision ranges) is derived in Python by the reference
(python/gpu_planning/substrate/robot_model_loader.py:32-115); no closed form
pins its exact radii, so this script imports that module by file path (it
depends only on PyYAML) and records, for each shipped robot, the derived
arrays it returns. Run here, where the reference is mounted:

    python tests/golden/make_loader_golden.py /root/reference

The fixture holds data only (inputs are the committed YAML assets).
"""
import hashlib
import importlib.util
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ASSETS = HERE.parents[1] / "prismplan_amd" / "assets" / "robots"


def main(ref_root: str) -> None:
    mod_path = Path(ref_root) / "code/gpu_planning/python/gpu_planning/substrate/robot_model_loader.py"
    spec = importlib.util.spec_from_file_location("ref_robot_model_loader", mod_path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {}
    for name in ("panda", "fetch", "baxter"):
        path = ASSETS / f"{name}.yaml"
        a = mod.load_robot_asset(str(path))
        out[name] = {
            "yaml_sha256": hashlib.sha256(path.read_bytes()).hexdigest(),
            "approx_spheres_flat": [float(v) for v in a["approx_spheres_flat"]],
            "approx_sphere_to_joint": [int(v) for v in a["approx_sphere_to_joint"]],
            "approx_self_cc_ranges_flat": [int(v) for v in a["approx_self_cc_ranges_flat"]],
            "n_t_slots": int(a["n_t_slots"]),
        }
    (HERE / "loader_golden.json").write_text(json.dumps(out, indent=1))
    print("wrote", HERE / "loader_golden.json")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
