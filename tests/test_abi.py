"""The C ABI: liblbm_amd.so loads and exports every function include/*.h
declares, the Python bindings cover exactly those, and the Python mirror of
LBMGrid has the C layout.  No GPU calls."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from lattice_fluid_dynamics_amd import lbm, model, superres, warmstart

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("lbm.h", "lbm_warmstart.h", "lbm_model.h", "lbm_superres.h")]
# lbm_model.h and lbm_superres.h keep the reference helpers' own names
MODEL_NAMES = (r"loadOBJ|freeModel|computeModelBounds|createTriangleBuffer|buildCollisionGrid|Voxelize_[A-Za-z]+|"
               r"SR_[A-Za-z]+")


def declared_functions(headers=HEADERS):
    names = set()
    for h in headers:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b((?:LBM|lbmnn)_[A-Za-z0-9_]+|" + MODEL_NAMES + r")\s*\(", text))
    return sorted(names)


def test_every_header_is_covered():
    assert sorted(HEADERS) == sorted(os.path.join(ROOT, "include", h)
                                     for h in os.listdir(os.path.join(ROOT, "include")) if h.endswith(".h"))


def test_header_declares_the_reference_api():
    names = set(declared_functions(HEADERS[:1]))
    reference_api = {  # simulation/lib/lbm.h:58-129
        "LBM_Create", "LBM_Free", "LBM_SetSolidAABB", "LBM_InitializeFlow", "LBM_Step",
        "LBM_GetVelocityBuffer", "LBM_ComputeDragForce", "LBM_ComputeDragForceDecomposed",
        "LBM_ComputeDragCoefficient", "LBM_ComputeLiftCoefficient", "LBM_SetSolidSphere",
        "LBM_AddGroundPlane", "LBM_ComputeProjectedArea", "LBM_SetSolidMesh"}
    assert reference_api <= names


def test_library_exports_every_declared_symbol():
    L = warmstart.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    bound = (set(lbm.SIGNATURES) | set(warmstart.WS_SIGNATURES) | set(model.SIGNATURES) | {"LBM_AutoPlaceModel"}
             | set(superres.SR_SIGNATURES))
    assert set(declared_functions()) == bound, set(declared_functions()) ^ bound


def test_grid_struct_layout_matches_c():
    fields = [f for f, _ in lbm.LBMGridStruct._fields_]
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "lbm.h"\nint main(void){\n'
    src += 'printf("%zu\\n", sizeof(LBMGrid));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof(LBMGrid, {f}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(lbm.LBMGridStruct)
    for f, off in zip(fields, out[1:]):
        assert getattr(lbm.LBMGridStruct, f).offset == off, f


def test_particle_record_is_32_bytes():
    src = '#include <stdio.h>\n#include "lbm.h"\nint main(void){printf("%zu", sizeof(Particle));return 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        assert subprocess.run([exe], capture_output=True, text=True, check=True).stdout == "32"


def test_model_records_match_c():
    """Model, CarBounds, GPUTriangle (48 bytes), ModelPlacement and VoxelGrid:
    the ctypes mirrors have the C sizes (simulation/lib/model_bounds.h:8-18)."""
    names = ["Model", "CarBounds", "GPUTriangle", "ModelPlacement", "VoxelGrid", "Vertex", "Face", "CollisionGrid"]
    src = '#include <stdio.h>\n#include "lbm_model.h"\nint main(void){\n'
    src += "".join(f'printf("%zu\\n", sizeof({n}));\n' for n in names) + "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    py = [ctypes.sizeof(getattr(model, n)) if hasattr(model, n) else 48 for n in names]
    assert out == py and out[2] == 48


def test_particle_flow_layout_matches_c():
    """LBMParticleFlow (include/lbm.h): size and every field offset of the
    ctypes mirror as the C compiler lays them out."""
    fields = [f for f, _ in lbm.LBMParticleFlow._fields_]
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "lbm.h"\nint main(void){\n'
    src += 'printf("%zu\\n", sizeof(LBMParticleFlow));\n'
    src += "".join(f'printf("%zu\\n", offsetof(LBMParticleFlow, {f}));\n' for f in fields) + "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(lbm.LBMParticleFlow)
    for f, off in zip(fields, out[1:]):
        assert getattr(lbm.LBMParticleFlow, f).offset == off, f


def test_create_without_device_fails_loudly(monkeypatch):
    from tests.conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError):
        lbm.LBM_Create(16, 8, 8, 0.1)
