"""The ISO_C_BINDING interface module (dalton_amd/fortran/ccgpu_interface.f90)
compiled with the image's Fortran compiler and linked against libccgpu.so:
a Fortran driver makes the calls of Dalton's CC drivers through it
(CCRHSN, CCSD_DIIS, CCSD_ECCSD, the energy loop, CC_LHTR, CC_ETA, the L0
solve, CC_D1AO, (T), CC_RWPRE) with CCGPU_DRYRUN=1, so every argument
reaches the library and every operand span is checked without a GPU; a bad
argument reaches QUIT through the status check."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang")
                                  else None)
MODULE = os.path.join(ROOT, "dalton_amd", "fortran", "ccgpu_interface.f90")
DRIVER = os.path.join(ROOT, "tests", "fortran", "dryrun_driver.f90")
LIBDIR = os.path.join(ROOT, "dalton_amd")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if FLANG is None:
        pytest.skip("no Fortran compiler in this image")
    if not os.path.exists(os.path.join(LIBDIR, "libccgpu.so")):
        pytest.skip("libccgpu.so is not built")
    d = tmp_path_factory.mktemp("fortran")
    exe = str(d / "dryrun_driver")
    subprocess.run([FLANG, "-c", MODULE, "-module-dir", str(d), "-o", str(d / "ccgpu_interface.o")], check=True,
                   capture_output=True, cwd=d)
    subprocess.run([FLANG, "-I", str(d), DRIVER, str(d / "ccgpu_interface.o"), "-L", LIBDIR, "-lccgpu",
                    f"-Wl,-rpath,{LIBDIR}", "-o", exe], check=True, capture_output=True, cwd=d)
    return exe


def _run(exe, mode):
    env = dict(os.environ, CCGPU_DRYRUN="1")
    return subprocess.run([exe, mode], capture_output=True, text=True, env=env, timeout=120)


def test_fortran_module_drives_every_wrapper(driver):
    r = _run(driver, "ok")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "dryrun driver ok" in r.stdout


def test_fortran_status_reaches_quit(driver):
    r = _run(driver, "bad")
    assert r.returncode != 0
    assert "QUIT: CC_LHTR:" in r.stdout, r.stdout + r.stderr
