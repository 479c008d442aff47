"""The C ABI boundary without a GPU: the HIP library loads, exports every function
include/vsg.h declares, the header compiles as C and C++, and the synthetic frame
generator is deterministic."""
import os
import re
import subprocess

import numpy as np

from vsg_amd import synth, vsg

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "vsg.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vsg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_api():
    names = _declared_functions()
    for required in ("vsg_scene_create", "vsg_scene_destroy", "vsg_scene_reset", "vsg_integrate",
                     "vsg_render_create", "vsg_render_frame", "vsg_render_download", "vsg_export_blocks",
                     "vsg_export_voxels", "vsg_heap_free_count", "vsg_timing_read"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = vsg.lib()
    names = _declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(vsg.exported_symbols()) == names  # the ctypes table covers the header exactly
    nm = subprocess.run(["nm", "-D", "--defined-only", vsg.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\bT (vsg_[a-z0-9_]+)", nm.stdout))
    assert set(names) <= exported


def test_library_has_gfx950_code_object():
    # the fat binary embedded in the library names its offload target
    data = open(vsg.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "use.c"
    src.write_text('#include "vsg.h"\nint main(void){ vsg_params p = {0}; (void)p; return vsg_version() > 0 ? 0 : 1; }\n')
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-c", "-I", os.path.join(REPO, "include"), str(src),
                    "-o", str(tmp_path / "use_c.o")], check=True)
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++", "-c", "-I", os.path.join(REPO, "include"),
                    str(src), "-o", str(tmp_path / "use_cpp.o")], check=True)


def test_struct_layouts_match_header():
    import ctypes as C
    assert C.sizeof(vsg.vsg_params) == 32
    assert C.sizeof(vsg.vsg_intrinsics) == 16
    assert C.sizeof(vsg.vsg_frame) == 32
    assert C.sizeof(vsg.vsg_counters) == 64


def test_params_mirror_reference_fields():
    p = vsg.Params(200000, 1 << 17, 0.004, 0.02)
    assert p.totalHashSize == 2000000 and p.totalBlockSize == (1 << 17) * 512
    assert p.SDFBlockSize == 8 and p.slotSize == 10


def test_synth_frames_deterministic():
    sc = synth.SynthScene(1)
    a, b = sc.render(7), synth.SynthScene(1).render(7)
    for k in ("points", "colors", "normals", "depth", "pose"):
        assert np.array_equal(a[k], b[k])
    d = a["depth"]
    assert 0.5 < (d > 0).mean() <= 1.0 and d[d > 0].min() >= 0.2 and d.max() <= 4.0
    # pointmap back-projects the depth at integer pixel coordinates
    u = np.arange(sc.W, dtype=np.float32)[None, :]
    x = (u - np.float32(sc.cx)) * d / np.float32(sc.fx)
    assert np.array_equal(a["points"][..., 0], np.where(d > 0, x, 0).astype(np.float32))
    # normals are unit length where set; zero on the border
    n = np.linalg.norm(a["normals"], axis=-1)
    assert np.allclose(n[n > 0], 1.0, atol=1e-5) and np.all(n[0] == 0) and np.all(n[:, -1] == 0)


def test_synth_trajectory_limits():
    sc = synth.SynthScene(1)
    P = [sc.pose(i) for i in range(sc.nframes)]
    for a, b in zip(P[:-1], P[1:]):
        assert np.linalg.norm(b[:3, 3] - a[:3, 3]) <= 0.05 + 1e-9
        R = a[:3, :3].T @ b[:3, :3]
        ang = np.degrees(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1)))
        assert ang <= 3.0
        assert np.allclose(b[:3, :3].T @ b[:3, :3], np.eye(3), atol=1e-12)
