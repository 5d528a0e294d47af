"""The C-ABI library loads and exports every function include/*.h declares; argument checks
that need no GPU return the documented error codes."""
from __future__ import annotations

import ctypes
import subprocess
import re
from pathlib import Path

import numpy as np

import arcadia_amd as A
from arcadia_amd import native

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(arc_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = _declared_functions()
    assert "arc_trace" in names and "arc_world_upload" in names and "arc_world_generate" in names
    lib = ctypes.CDLL(str(native.LIB_PATH))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert missing == []
    assert names <= set(native.EXPORTED) | {n for n in names if hasattr(lib, n)}


def test_header_struct_sizes_match_bindings(tmp_path):
    """sizeof/offsetof from the C compiler against the Python bindings."""
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "arcadia_hip.h"\n#include "arcadia_world.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(arc_node64), "
        "sizeof(arc_brick64), sizeof(arc_ray), sizeof(arc_hit), sizeof(arc_trace_result), "
        "sizeof(arc_trace_params), sizeof(arc_gen_params), sizeof(arc_camera), sizeof(arc_gpu_patch), "
        "offsetof(arc_hit, steps), sizeof(arc_gbuffer_params), offsetof(arc_gbuffer_params, width), "
        "sizeof(arc_shadow_params), sizeof(arc_shadow_rec), offsetof(arc_shadow_rec, outcome), "
        "sizeof(arc_rigid_body), offsetof(arc_rigid_body, world_sphere), sizeof(arc_rigid_hit), "
        "offsetof(arc_rigid_hit, body), sizeof(arc_gi_reservoir), sizeof(arc_di_reservoir), "
        "sizeof(arc_emissive_source), sizeof(arc_alias_entry), sizeof(arc_restir_params));return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [native.NODE_DTYPE.itemsize, native.BRICK_DTYPE.itemsize, native.RAY_DTYPE.itemsize,
            native.HIT_DTYPE.itemsize, native.RESULT_DTYPE.itemsize, ctypes.sizeof(native.TraceParams),
            ctypes.sizeof(native.GenParams), ctypes.sizeof(native.Camera), ctypes.sizeof(native.GpuPatch),
            native.HIT_DTYPE.fields["steps"][1], ctypes.sizeof(native.GbufferParams),
            native.GbufferParams.width.offset, ctypes.sizeof(native.ShadowParams), native.SHADOW_REC_DTYPE.itemsize,
            native.SHADOW_REC_DTYPE.fields["outcome"][1], native.RIGID_BODY_DTYPE.itemsize,
            native.RIGID_BODY_DTYPE.fields["world_sphere"][1], native.RIGID_HIT_DTYPE.itemsize,
            native.RIGID_HIT_DTYPE.fields["body"][1], native.GI_RESERVOIR_DTYPE.itemsize,
            native.DI_RESERVOIR_DTYPE.itemsize, native.EMISSIVE_DTYPE.itemsize, native.ALIAS_DTYPE.itemsize,
            ctypes.sizeof(native.RestirParams)]
    assert got == want


def test_host_argument_errors():
    lib = A.lib()
    assert lib.arc_world_host_create(None, None) == native.ARC_ERR_INVALID
    g = (ctypes.c_uint32 * 3)(0, 1, 1)
    h = ctypes.c_void_p()
    assert lib.arc_world_host_create(g, ctypes.byref(h)) == native.ARC_ERR_INVALID
    assert lib.arc_create(0, None) == native.ARC_ERR_INVALID
    bad = A.gen_params(A.ARC_GEN_HEIGHTMAP)
    bad.mat_min = 0
    w = A.HostWorld((1, 1, 1))
    assert lib.arc_world_generate(w._h, ctypes.byref(bad), 1) == native.ARC_ERR_INVALID
    assert lib.arc_world_host_pools(w._h, None, None, None, None, None, None) == native.ARC_ERR_INVALID  # not packed


def test_ray_generation_is_deterministic_and_sane():
    cam = A.camera_look((8.1, 14.2, 7.9), 0.4, -0.6, 90.0, 64, 36)
    r1, pix = A.primary_rays(cam, 64, 36)
    r2, _ = A.primary_rays(cam, 64, 36)
    assert r1.tobytes() == r2.tobytes()
    assert sorted(pix.tolist()) == list(range(64 * 36))
    d = np.stack([r1["dx"], r1["dy"], r1["dz"]], 1).astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    fwd = np.array(cam.forward[:], np.float64)
    assert (d @ fwd).min() > 0.4  # corner of a 90 deg x 16:9 frustum: cos = 1/sqrt(1 + 1 + (16/9)^2)
    # first 64 rays are the top-left 8x8 tile
    assert set(pix[:64].tolist()) == {y * 64 + x for y in range(8) for x in range(8)}


def test_bounce_rays_leave_the_surface_on_the_normal_side():
    n = 1000
    rng = np.random.default_rng(0)
    pr = np.zeros(n, A.RAY_DTYPE)
    pr["dy"] = -1.0
    pr["ox"], pr["oy"], pr["oz"] = rng.uniform(1, 15, n), 15.0, rng.uniform(1, 15, n)
    hits = np.zeros(n, A.HIT_DTYPE)
    hits["flags"] = 1
    hits["t_voxels"] = 100.0
    hits["face"] = rng.integers(0, 6, n)
    br, src = A.bounce_rays(pr, hits, seed=3)
    assert len(br) == n and np.array_equal(src, np.arange(n))
    for i in range(n):
        f = int(hits["face"][i])
        nrm = np.zeros(3)
        nrm[f >> 1] = -1.0 if f & 1 else 1.0
        d = np.array([br["dx"][i], br["dy"][i], br["dz"][i]])
        lifted = nrm + np.array([0, 0.22, 0])
        assert d @ lifted > -1e-6
        hp = np.array([pr["ox"][i], pr["oy"][i] - 100.0 / 16.0, pr["oz"][i]])
        o = np.array([br["ox"][i], br["oy"][i], br["oz"][i]])
        assert np.allclose(o - hp, 0.10 * nrm, atol=1e-5)
