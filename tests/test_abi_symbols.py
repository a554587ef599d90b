"""The C-ABI library loads and exports every function include/*.h declares.

No compute call is made (no device is needed): only symbol resolution and
the error path of a create without a device.
"""
import ctypes
import glob
import os
import re

import fluid_amd as fa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:extern\s+)?(?:const\s+)?(?:struct\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*"
                  r"(fluid_[a-z0-9_]+)\s*\(", re.M)


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(DECL.findall(text))
    return names


def test_headers_declare_the_step_surface():
    names = declared()
    for required in ("fluid_sim_create", "fluid_sim_substeps", "fluid_sim_frame", "fluid_sim_read_stats",
                     "fluid_sim_tracers", "fluid_pack_sim_params", "fluid_pack_step", "fluid_scan_u32"):
        assert required in names


def test_library_exports_every_declared_function():
    L = ctypes.CDLL(fa.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(L, n)]
    assert not missing, missing


def test_create_fails_cleanly_without_a_device_or_params():
    L = fa.lib()
    assert not L.fluid_sim_create(None, 0, 0)
    assert b"SimParams" in L.fluid_sim_last_error() or b"device" in L.fluid_sim_last_error()
