"""The C ABI without a GPU: libddns.so loads, exports every entry point that
include/ddns.h declares, its structs have the same layout in C, ctypes and Fortran,
the defaults agree, and the parameter.d reader reads the reference's files (these
calls touch no device)."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ddns.h"
GOLD = Path(__file__).parent / "golden"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ddns_[a-z0-9_]+)\s*\(", text)))


def test_header_and_ctypes_table_agree():
    from mpi_2d_domain_decomp_turb_particles_amd import _abi
    assert set(_declared()) == set(_abi.SIGNATURES)


def test_library_exports_every_declared_symbol(ddns):
    lib = ddns.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(ddns.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (ddns_[a-z0-9_]+)$", out, flags=re.M))
    assert set(_declared()) <= exported


def _c_layout(struct, fields):
    """offsetof / sizeof of a struct from include/ddns.h as gcc lays it out."""
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"ddns.h\"\nint main(void){\n"
    src += f'printf("%zu\\n", sizeof({struct}));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof({struct}, {f}));\n'
    src += "return 0;}\n"
    return src


@pytest.mark.parametrize("cname,pyname", [("ddns_params", "Params"), ("ddns_collision_counts", "CollisionCounts"),
                                           ("ddns_timing", "Timing")])
def test_struct_layout_c_vs_ctypes(tmp_path, cname, pyname):
    from mpi_2d_domain_decomp_turb_particles_amd import _abi
    cls = getattr(_abi, pyname)
    names = [f[0] for f in cls._fields_]
    (tmp_path / "l.c").write_text(_c_layout(cname, names))
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(tmp_path / "l.c"), "-o", str(tmp_path / "l")], check=True)
    vals = [int(x) for x in subprocess.run([str(tmp_path / "l")], capture_output=True, text=True).stdout.split()]
    assert vals[0] == C.sizeof(cls)
    for name, off in zip(names, vals[1:]):
        assert getattr(cls, name).offset == off, name


def test_fortran_bind_c_types_match(tmp_path):
    """The bind(C) derived types of fortran/ddns_mod.f90 have the C sizes."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not Path(fc).exists():
        pytest.skip("no Fortran compiler")
    from mpi_2d_domain_decomp_turb_particles_amd import _abi
    prog = tmp_path / "t.f90"
    prog.write_text("""program t
  use, intrinsic :: iso_c_binding
  use ddns_mod
  type(ddns_params) :: p
  type(ddns_collision_counts) :: c
  type(ddns_timing) :: m
  print *, c_sizeof(p), c_sizeof(c), c_sizeof(m)
end program t
""")
    mod = ROOT / "mpi_2d_domain_decomp_turb_particles_amd" / "fortran" / "ddns_mod.f90"
    pkg = ROOT / "mpi_2d_domain_decomp_turb_particles_amd"
    if not (pkg / "libddns.so").exists():
        pytest.skip("libddns.so not built")
    subprocess.run([fc, "-J", str(tmp_path), str(mod), str(prog), "-o", str(tmp_path / "t"), f"-L{pkg}", "-lddns",
                    f"-Wl,-rpath,{pkg}"], check=True, cwd=tmp_path)
    got = [int(x) for x in subprocess.run([str(tmp_path / "t")], capture_output=True, text=True).stdout.split()]
    assert got == [C.sizeof(_abi.Params), C.sizeof(_abi.CollisionCounts), C.sizeof(_abi.Timing)]


def test_defaults_c_equal_python(ddns):
    lib = ddns.load_library()
    pc = ddns.Params()
    assert lib.ddns_params_default(C.byref(pc)) == 0
    pp = ddns.default_params()
    for name, _ in ddns.Params._fields_:
        a, b = getattr(pc, name), getattr(pp, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name


def test_parameter_file_256_det(ddns):
    p = ddns.read_parameter_file(GOLD / "parameter.d_256_det")
    assert (p.nx, p.ny, p.nz, p.nprocY, p.nstep) == (256, 256, 256, 8, 100000)
    assert (p.dt, p.rnu, p.time) == (0.001, 0.002, 0.0)
    assert (p.new_flow, p.iseed_m0, p.forcing) == (1, 123, 1)
    assert list(p.force) == [1.242477, 0.391356]
    assert (p.iseedf, p.ieout, p.ispec, p.divstat, p.st) == (232300, 10, 1000, 1, 2.0)
    assert (p.etk_DNS, p.tauk_DNS) == (0.01392, 0.09694)
    assert (p.ediss_CLOUD, p.viscosity_CLOUD, p.gravity) == (400.0, 0.17, 980.0)
    assert (p.new_particle, p.nonoverlap, p.hdi_included, p.hdi_trunc) == (1, 0, 0, 50.0)
    assert (p.npart, p.ndrag, p.nset) == (2000000, 0, 2)
    assert list(p.radius)[:2] == [15.0, 20.0]


def test_parameter_file_128_stoch(ddns):
    p = ddns.read_parameter_file(GOLD / "parameter.d_128_stoch")
    assert (p.nx, p.forcing, p.iseedf, p.npart, p.nset) == (128, 2, 232300, 500000, 2)
    assert (p.dt, p.rnu) == (0.0001, 0.09449408)
    assert list(p.radius)[:2] == [10.0, 15.0]


def test_parameter_file_hdi_forces_nonoverlap(ddns):
    p = ddns.read_parameter_file(GOLD / "parameter.d_64_short")
    assert p.hdi_included == 1 and p.nonoverlap == 1          # turb.f90:171
    assert (p.nx, p.nstep, p.npart, p.new_flow) == (64, 5, 20000, 1)


def test_missing_parameter_file_is_an_io_error(ddns):
    with pytest.raises(ddns.DDNSError) as e:
        ddns.read_parameter_file(GOLD / "no_such_file")
    assert e.value.code == ddns.DDNS_ERR_IO


def test_missing_library_fails_loudly(tmp_path):
    import mpi_2d_domain_decomp_turb_particles_amd as dd
    with pytest.raises(ImportError):
        dd.load_library(tmp_path / "libddns.so")


def test_fortran_arrays_only(ddns):
    with pytest.raises(TypeError):
        ddns._dp(np.zeros((4, 4)))          # C order is refused
