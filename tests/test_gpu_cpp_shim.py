"""The reference's C++ class API (include/vsg_voxelscene.hpp) driven like the
reference's voxelSG.cpp loop, from a C++ executable, gives the same scene and
renders as the Python mirror over the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

from fusion_util import GpuRun
from vsg_amd import synth

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(synth.__file__), "lib", "vsg_shim_example")


@pytest.mark.parametrize("workload,frames", [(0, 6), (1, 3)])
def test_cpp_shim_matches_python_path(torch_cuda, workload, frames):
    out = subprocess.run([EXE, str(workload), str(frames)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    tok = out.stdout.split()
    got = {tok[i]: tok[i + 1] for i in range(0, len(tok) - 1, 2)}
    wl = synth.WORKLOADS[workload]
    sc = synth.SynthScene(workload)
    g = GpuRun(torch_cuda, wl)
    for i in range(frames):
        f = sc.render(i)
        g.integrate(f, sc.intrinsics)
    r = g.raycast(f, sc.intrinsics)
    d = r["depth"]
    k = np.arange(d.size)
    ok = d < 1e9
    checksum = float(np.sum(d[ok].astype(np.float64) * ((k[ok] % 97) + 1)))
    assert int(got["frames"]) == frames
    assert int(got["blocks"]) == g.scene.counters()["num_allocated"]
    assert (int(got["hits"]), int(got["misses"])) == (r["hits"], r["misses"])
    assert abs(float(got["checksum"]) - checksum) <= 1e-9 * abs(checksum)
