"""Writes reference_materials.json: the principled parameters of every
material asset the reference ships (materials/*.pbrdf), read with this
package's .pbrdf reader (gi_pbrdf_read, which the scene-ingestion tests pin
against an independent restatement) starting from a white Lambert record.
The assets are data; the fixture lets GPU tests render the reference's own
materials where the reference tree is absent.

    python tests/golden/make_reference_materials.py [/root/reference]
"""
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from globillum_amd import scene_io, scenes  # noqa: E402

FIELDS = ["color", "metallic", "subsurface", "specular", "roughness", "specular_tint", "anisotropic", "sheen",
          "sheen_tint", "clearcoat", "clearcoat_gloss", "opacity", "transmission"]


def main(root: str) -> None:
    out = {}
    for path in sorted(glob.glob(os.path.join(root, "materials", "*.pbrdf"))):
        m = scene_io.read_pbrdf(path, scenes.lambert((1.0, 1.0, 1.0)))
        rec = {}
        for f in FIELDS:
            v = m[f][0]
            rec[f] = [float(x) for x in v] if getattr(v, "shape", ()) else float(v)
        out[os.path.splitext(os.path.basename(path))[0]] = rec
    with open(os.path.join(HERE, "reference_materials.json"), "w") as fh:
        json.dump({"source": "materials/*.pbrdf of the reference", "materials": out}, fh, indent=1, sort_keys=True)
    print(f"{len(out)} materials")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
