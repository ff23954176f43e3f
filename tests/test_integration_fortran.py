"""The Fortran binding documented for maintainers compiles and links.

The ```fortran blocks of INTEGRATION.md are extracted and compiled with ROCm's
flang: the iso_c_binding module on its own; the drop-in iccg against the
reference's own module files (types, parameters, geometry, sparse_matrix,
title_mod) where `make -C oracle ref` has built them; and a program that
calls fc_init, linked against libfcap.so (linked, not run: no GPU here).
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
FLANG = "/opt/rocm/lib/llvm/bin/flang"
LIB_DIR = os.path.join(ROOT, "freecappuccino_amd")
REF_MODS = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="ROCm flang not installed")


def _blocks():
    with open(os.path.join(ROOT, "INTEGRATION.md")) as f:
        text = f.read()
    return re.findall(r"```fortran\n(.*?)```", text, re.S)


def _run(cmd, cwd):
    p = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert p.returncode == 0, f"{' '.join(cmd)}\n{p.stdout}\n{p.stderr}"


def test_binding_module_compiles_and_links(tmp_path):
    blocks = _blocks()
    assert len(blocks) >= 2
    binding = next(b for b in blocks if b.lstrip().startswith("module fcap_binding"))
    (tmp_path / "fcap_binding.f90").write_text(binding)
    _run([FLANG, "-c", "fcap_binding.f90"], tmp_path)
    assert (tmp_path / "fcap_binding.mod").exists()
    # a program that starts the library: resolves fc_init and fc_set_report in libfcap.so
    (tmp_path / "main.f90").write_text("program p\n  use fcap_binding\n  call fcap_start()\nend program\n")
    if not os.path.exists(os.path.join(LIB_DIR, "libfcap.so")):
        pytest.skip("libfcap.so not built")
    _run([FLANG, "main.f90", "fcap_binding.o", f"-L{LIB_DIR}", "-lfcap", f"-Wl,-rpath,{LIB_DIR}", "-o", "main"],
         tmp_path)
    assert (tmp_path / "main").exists()


def test_drop_in_solver_compiles_against_the_reference_modules(tmp_path):
    if not os.path.exists(os.path.join(REF_MODS, "sparse_matrix.mod")):
        pytest.skip("reference module files not built (make -C oracle ref)")
    blocks = _blocks()
    binding = next(b for b in blocks if b.lstrip().startswith("module fcap_binding"))
    solver = next(b for b in blocks if b.lstrip().startswith("subroutine iccg"))
    (tmp_path / "fcap_binding.f90").write_text(binding)
    (tmp_path / "fcap_solvers.f90").write_text(solver)
    _run([FLANG, "-c", "fcap_binding.f90"], tmp_path)
    _run([FLANG, "-c", "-I", REF_MODS, "-I", str(tmp_path), "fcap_solvers.f90"], tmp_path)
    assert (tmp_path / "fcap_solvers.o").exists()
