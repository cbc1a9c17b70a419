"""The C-ABI library builds for gfx950, loads, and exports exactly what
include/meso_hip.h declares.  No compute call is made (no GPU here)."""
import ctypes as C
import os
import re

from mesoengine_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "meso_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[a-z_][\w \*]*?\b(meso_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_binding_set():
    fns = header_functions()
    assert len(fns) > 50
    assert fns == L.EXPORTED, set(fns) ^ set(L.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(L.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing
    L.load()


def test_library_is_built_for_gfx950():
    data = open(L.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_abi_version_and_error_paths_without_gpu():
    lib = L.load()
    assert lib.meso_abi_version() == 1
    h = C.c_void_p()
    st = lib.meso_context_create(0, C.byref(h))
    if st != L.OK:  # no GPU in this container: a clean error, no crash
        assert st == L.ERR_RUNTIME or st == L.ERR_ARG
        assert lib.meso_last_error_message()
    else:
        lib.meso_context_destroy(h)
    assert lib.meso_buffer_create(None, 16, None, C.byref(h)) == L.ERR_ARG
    assert b"NULL" in lib.meso_last_error_message()
    assert lib.meso_buffer_device_address(None, 0) == 0
    assert lib.meso_mesh_workspace_bytes(4096) == 256 + 8 * 4096
    assert lib.meso_raymarch(None, None, None, 0, None, None, None) == L.ERR_ARG
    assert lib.meso_serve_submit(None, None, None, 0, None, None) == L.ERR_ARG


def test_struct_sizes_match_header_layout():
    # the ctypes mirrors carry the same sizes the C compiler gives the structs
    assert C.sizeof(L.Camera) == 4 + 9 * 4 + 4 + 8
    assert C.sizeof(L.RayOutputs) == 56
    assert C.sizeof(L.ChunkSource) == 32
    assert C.sizeof(L.GenerateDesc) == 32


def test_engine_instance_builds_and_rejects_bad_arguments():
    """The headless C++ instance links against the library; without a GPU
    it fails with the library's message instead of crashing."""
    import subprocess
    b = os.path.join(ROOT, "mesoengine_amd", "_build")
    exe = os.path.join(b, "meso_instance")
    assert os.path.exists(exe) and os.path.exists(os.path.join(b, "libmeso_engine.so"))
    p = subprocess.run([exe, "--scene", "nowhere"], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr
    eng = C.CDLL(os.path.join(b, "libmeso_engine.so"))
    assert eng is not None
    sb = os.path.join(b, "meso_serve_bench")
    p = subprocess.run([sb], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2 and "usage" in p.stderr
    p = subprocess.run([sb, os.path.join(b, "no_such_requests.bin")], capture_output=True,
                       text=True, timeout=60)
    assert p.returncode == 1 and "cannot read" in p.stderr


def test_byref_keeps_a_temporary_struct_alive():
    # byref(make_struct()) is passed straight into calls: the pointer object
    # must hold the struct, or the allocator may reuse its bytes (the first
    # 8 of a freed block hold the free list) before the library reads them
    p = L.byref(L.Shading(C.sizeof(L.Shading), (C.c_float * 3)(0.4, 1.0, 0.3), 0, 4, 16.0))
    held = p._obj
    assert isinstance(held, L.Shading)
    assert list(held.light) == [C.c_float(0.4).value, 1.0, C.c_float(0.3).value]
    assert C.addressof(held) == C.cast(p, C.c_void_p).value
