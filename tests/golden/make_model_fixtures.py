"""Writes tests/golden/model_meshes.npz from the reference's OBJ assets.

For each of simulation/assets/3d-files/{car-model, ahmed_25deg_m,
ahmed_35deg_m}.obj it stores
  <name>_tris   (n, 12) float32  the placed triangles LBM_SetSolidMesh takes
  <name>_bbox   (6,) float32     their bounds (minX minY minZ maxX maxY maxZ)
  <name>_place  (5,) float32     scale, offsetX, offsetY, offsetZ, rotationY
  <name>_vox64  packed bits      Voxelize_FromModel(R=64, padding 0.1,
                                 align 1, solid) as computed by the
                                 reference's own voxelize.c (oracle/_ref)
The triangles come from this package's OBJ reader and placement; the
voxel grid comes from the reference build, so the fixture pins both.
Run in a container holding the reference tree, after `make -C oracle ref`:
    python tests/golden/make_model_fixtures.py
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from lattice_fluid_dynamics_amd import model as M  # noqa: E402

ASSETS = "/root/reference/simulation/assets/3d-files"
MESHES = {"car": "car-model.obj", "ahmed25": "ahmed_25deg_m.obj", "ahmed35": "ahmed_35deg_m.obj"}


def main():
    ref = M.bind(ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_model.so")), placement=False)
    out = {}
    for key, fname in MESHES.items():
        path = os.path.join(ASSETS, fname)
        tris, bbox, place = M.mesh_from_obj(path)
        rm = M.LoadedModel(path, lib=ref)
        ref_tris = rm.triangles(*place)
        assert np.array_equal(ref_tris.view(np.uint32), tris.view(np.uint32)), key
        vox = rm.voxelize(64, 0.1, 1, M.VOXELIZE_MODE_SOLID)
        rm.close()
        out[key + "_tris"] = tris
        out[key + "_bbox"] = np.array(bbox, np.float32)
        out[key + "_place"] = np.array(place, np.float32)
        out[key + "_vox64"] = np.packbits(vox.reshape(-1))
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "model_meshes.npz"), **out)


if __name__ == "__main__":
    main()
