"""The C ABI library loads and exports every entry point of include/fd4d.h
(CPU only: no kernel is launched)."""
import ctypes
import os
import re
import subprocess

import pytest


def _declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points(fd):
    names = _declared(fd.HEADER)
    for n in ["fd_init", "fd_destroy", "fd_set_backface_cull", "fd_add_model", "fd_set_sample_buffers",
              "fd_set_output_buffers", "fd_register_res", "fd_unregister_res", "fd_set_bin_num", "fd_set_view",
              "fd_set_scenes_esp", "fd_set_scene_params", "fd_set_area_light", "fd_launch", "fd_move_model"]:
        assert n in names


def test_library_exports_every_declared_symbol(fd):
    lib = fd.lib()
    names = _declared(fd.HEADER)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", fd.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    assert set(names) <= exported


def test_host_only_calls(fd):
    lib = fd.lib()
    assert lib.fd_mask_words(16) == 16 and lib.fd_mask_words(32) == 32 and lib.fd_mask_words(33) == 66
    assert lib.fd_mask_words(0) == fd.FD_ERR_ARG and lib.fd_mask_words(65) == fd.FD_ERR_ARG
    assert lib.fd_status_string(fd.FD_ERR_CAPACITY) == b"pair arena too small"
    # null context is rejected, never dereferenced
    assert lib.fd_set_view(None, 4, 4) == fd.FD_ERR_ARG
    assert lib.fd_launch(None, None, None) == fd.FD_ERR_ARG
    assert lib.fd_init(None) == fd.FD_ERR_ARG


def test_fd_namespace_mirror_links(fd):
    """include/fd_api.hpp (the FD:: mirror) compiles against the C ABI and its
    driver program is built and linked to libfd4d.so (no GPU call)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = os.path.join(root, "include")
    src = '#include "fd_api.hpp"\nint main() { FD::AreaLightDes d; (void)d; return FD::FDsetBinNum(16) == 0; }\n'
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, "-x", "c++", "-"],
                       input=src, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    exe = os.path.join(root, "4dzb_amd", "fd_api_check")
    assert os.path.exists(exe)
    out = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    for n in ["fd_init", "fd_launch", "fd_set_area_light", "fd_add_model"]:
        assert n in out


def test_drivers_compile_and_bench_refuses_several_gpus():
    """bench.py and __graft_entry__.py are valid Python; bench.py --gpus 2
    exits with a message before it touches a device (one GPU per run)."""
    import py_compile
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for f in ("bench.py", "__graft_entry__.py"):
        py_compile.compile(os.path.join(root, f), doraise=True)
    r = subprocess.run(["python3", os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300, cwd=root)
    assert r.returncode != 0 and "--gpus 1" in r.stderr
