"""MI355X-native VoxelSG depth-fusion path (integrate + raycast).

Modules:
    vsg    ctypes host mirror of the reference's VoxelScene / RayCastRender API
           over the C ABI in include/vsg.h (lib/libvsg.so, HIP kernels for gfx950)
    synth  seeded synthetic depth frames of the fusion workloads (lib/libvsg_synth.so)
    autodiff  torch.autograd wrapper of the raycast adjoint (rendered depth as a
              function of a pose perturbation and of the stored SDF)

The directory name is not a Python identifier; load it with :func:`load` from a
file path, e.g. ``importlib`` or the helper in ``__graft_entry__``.
"""
import importlib.util
import os
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
PKG_NAME = "vsg_amd"


def load():
    """Returns the package registered as ``vsg_amd`` (idempotent)."""
    if PKG_NAME in sys.modules:
        return sys.modules[PKG_NAME]
    spec = importlib.util.spec_from_file_location(
        PKG_NAME, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[PKG_NAME] = mod
    spec.loader.exec_module(mod)
    return mod
