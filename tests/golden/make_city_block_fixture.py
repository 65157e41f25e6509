"""Writes tests/golden/city_block_gltf.npz: the reference's own city-block
scene (assets/scenes/city-block.gltf + .bin of the reference repository) as
loaded by irtt.gltf — global vertex / triangle arrays, per-triangle
materials, per-vertex initial and fixed temperatures, object ranges and the
sky patches — so GPU tests can run the reference's scene where the
reference repository is absent.

    python tests/golden/make_city_block_fixture.py [/path/to/reference]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "inverse-radiative-thermal-transport_amd"))

from irtt import gltf  # noqa: E402

FIXTURE = os.path.join(HERE, "city_block_gltf.npz")


def scene_arrays(sd):
    sky = sd.meta["sky"]
    objs = [{k: v for k, v in o.items()} for o in sd.meta["objects"]]
    return {"xyz": sd.xyz, "idx": sd.idx, "rho_d": sd.rho_d, "rho_s": sd.rho_s,
            "thickness": sd.thickness, "density": sd.density,
            "heat_capacity": sd.heat_capacity, "conductivity": sd.conductivity,
            "flags": sd.flags, "kelvin": sd.kelvin, "fixed": sd.fixed,
            "sky_kelvin": sky["kelvin"], "sky_patch_vertices": sky["patch_vertices"],
            "sky_values": sky["values"], "sky_hours": np.array(sky["hours"]),
            "objects": np.array(json.dumps(objs))}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    sd = gltf.load(os.path.join(ref, "assets", "scenes", "city-block.gltf"))
    np.savez_compressed(FIXTURE, **scene_arrays(sd))
    print(f"wrote {FIXTURE}: {sd.nv} vertices, {sd.nt} triangles")


if __name__ == "__main__":
    main()
