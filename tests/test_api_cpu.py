"""C-ABI and host-logic tests that need no GPU: the library loads, exports every symbol
declared in include/parament.h, and the error paths of the reference's test_error.py
(which the host code resolves before any device work) behave the same."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import equiprop_oracle as eo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "parament.h")).read()
    return sorted(set(re.findall(r"^LIBSPEC\s+[\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from parament_amd import paramentlib
    names = _declared_symbols()
    assert len(names) >= 29
    missing = [n for n in names if not hasattr(paramentlib.lib, n)]
    assert not missing, missing


def test_header_compiles_as_c_and_a_c_program_links(tmp_path):
    """include/parament.h is a C99 header: a C program that takes the address of every declared
    entry point compiles with gcc -Wall -Werror, links against libparament.so, and runs the
    host-only calls (no device work)."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    names = _declared_symbols()
    src = tmp_path / "abi.c"
    src.write_text(
        "#include <stdio.h>\n#include <string.h>\n#include \"parament.h\"\n"
        "typedef void (*entry)(void);\n"
        "static entry entries[] = {\n" + ",\n".join(f"    (entry){n}" for n in names) + "\n};\n"
        "int main(void) {\n"
        "    Parament_Context_f32 *h = NULL;\n"
        "    if (Parament_create(&h) != PARAMENT_STATUS_SUCCESS || !h) return 1;\n"
        "    if (Parament_destroy(h) != PARAMENT_STATUS_SUCCESS) return 2;\n"
        "    if (strcmp(Parament_errorMessage(PARAMENT_STATUS_SUCCESS), \"Success\") != 0) return 3;\n"
        "    if (Parament_selectIterationCycles_fp32(1.0, 0.01) < 3) return 4;\n"
        "    printf(\"%d entries\\n\", (int)(sizeof entries / sizeof entries[0]));\n"
        "    return entries[0] ? 0 : 5;\n"
        "}\n")
    libdir = os.path.join(ROOT, "parament_amd")
    exe = tmp_path / "abi"
    cmd = ["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", libdir, "-lparament", f"-Wl,-rpath,{libdir}"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert run.stdout.strip() == f"{len(names)} entries"


def test_reference_symbols_are_declared():
    # the 17 reference exports (parament.h:155-400) + device_info + OneNorm*
    ref = ["Parament_create", "Parament_destroy", "Parament_setHamiltonian", "Parament_equiprop",
           "Parament_selectIterationCycles_fp32", "Parament_setIterationCyclesManually",
           "Parament_automaticIterationCycles", "Parament_peekAtLastError", "Parament_errorMessage",
           "Parament_create_fp64", "Parament_destroy_fp64", "Parament_setHamiltonian_fp64",
           "Parament_equiprop_fp64", "Parament_setIterationCyclesManually_fp64",
           "Parament_automaticIterationCycles_fp64", "Parament_peekAtLastError_fp64",
           "Parament_selectIterationCycles_fp64", "device_info", "OneNorm", "OneNorm_fp64"]
    assert set(ref) <= set(_declared_symbols())


def test_iteration_tables_agree_with_oracle():
    from parament_amd import paramentlib
    lib = paramentlib.lib
    for x in np.r_[np.logspace(-6, 1.2, 300), [b for b, _ in eo.FP32_TABLE], [b for b, _ in eo.FP64_TABLE]]:
        assert lib.Parament_selectIterationCycles_fp32(float(x), 1.0) == eo.select_iteration_cycles_fp32(x, 1.0)
        assert lib.Parament_selectIterationCycles_fp64(float(x), 1.0) == eo.select_iteration_cycles_fp64(x, 1.0)


def test_one_norm_agrees_with_oracle():
    from parament_amd import paramentlib
    rng = np.random.default_rng(0)
    for d in (1, 3, 8):
        m = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
        assert abs(paramentlib.lib.OneNorm_fp64(np.ascontiguousarray(m.ravel()), d) - eo.one_norm(m)) < 1e-12
        m32 = m.astype(np.complex64)
        assert abs(paramentlib.lib.OneNorm(np.ascontiguousarray(m32.ravel()), d) - eo.one_norm(m32)) < 1e-5
    # max row sum of the row-major input, not column sums: 5 here, 7 for the transpose
    m = np.array([[1, 2j], [0, 3 + 4j]])
    for a, want in ((m, 5.0), (m.T, 7.0)):
        assert paramentlib.lib.OneNorm_fp64(np.ascontiguousarray(a.ravel()), 2) == want
        assert paramentlib.lib.OneNorm(np.ascontiguousarray(a.astype(np.complex64).ravel()), 2) == want


def test_error_messages_match_reference_strings():
    from parament_amd import paramentlib
    msg = lambda c: paramentlib.lib.Parament_errorMessage(c).decode()
    assert msg(0) == "Success"
    assert msg(50) == "Invalid value."
    assert msg(70) == "Timestep too large"
    assert msg(80) == "No hamiltonian set"
    assert msg(90) == "Invalid quadrature selection."
    assert msg(1000) == "Unknown error code"


# ---- test_error.py of the reference (test_error.py:26-54); these resolve on the host ----

def test_use_after_destruction():
    import parament_amd as parament
    context = parament.Parament()
    context.destroy()
    with pytest.raises(RuntimeError, match='Attempting to use a context that has been destroyed'):
        context.set_hamiltonian(np.eye(2), np.eye(2))


def test_propagate_no_hamiltonian():
    import parament_amd as parament
    with parament.Parament() as context:
        with pytest.raises(RuntimeError, match='No hamiltonian set'):
            context.equiprop(1.0)


def test_invalid_quadrature_selection():
    import parament_amd as parament
    with parament.Parament() as context:
        with pytest.raises(ValueError, match='Invalid quadrature selection'):
            context.set_hamiltonian(np.eye(2), np.eye(2), use_magnus=True, quadrature_mode='none')
        with pytest.raises(ValueError, match='Invalid quadrature selection'):
            context.set_hamiltonian(np.eye(2), np.eye(2), use_magnus=True, quadrature_mode='midpoint')


def test_unknown_quadrature_string():
    import parament_amd as parament
    with parament.Parament() as context:
        with pytest.raises(ValueError, match='unknown quadrature mode selected'):
            context.set_hamiltonian(np.eye(2), np.eye(2), quadrature_mode='trapezoid')


def test_c_level_no_hamiltonian_code():
    from parament_amd import paramentlib
    lib = paramentlib.lib
    h = ctypes.c_void_p()
    assert lib.Parament_create_fp64(ctypes.byref(h)) == 0
    out = np.zeros(4, np.complex128)
    assert lib.Parament_equiprop_fp64(h, np.zeros(2, np.complex128), 1.0, 2, 1, out) == 80
    assert lib.Parament_peekAtLastError_fp64(h) == 80
    assert lib.Parament_destroy_fp64(h) == 0
    assert lib.Parament_destroy_fp64(None) == 0


def test_manual_cycles_validation():
    import parament_amd as parament
    with parament.Parament() as context:
        for bad in (0, 1, 2, 4, 10, 49):
            with pytest.raises(ValueError, match='Invalid value'):
                context.set_iteration_cycles(bad)
        context.set_iteration_cycles(9)
        context.automatic_iteration_cycles()


def test_precision_argument_checks():
    import parament_amd as parament
    with pytest.raises(ValueError):
        parament.Parament('fp16')
    # complex64 contexts default to the mixed pipeline, complex128 contexts to FP64 throughout
    with parament.Parament() as context:
        assert context.precision_mode == 'mixed'
        context.set_precision_mode('native')
        assert context.precision_mode == 'native'
    with parament.Parament('fp64') as context:
        assert context.precision_mode == 'native'
        for mode in ('mixed', 'mixed_fp32', 'mixed_bf16x2', 'native'):
            context.set_precision_mode(mode)
            assert context.precision_mode == mode
        with pytest.raises(ValueError):
            context.set_precision_mode('bf16')
    # the C entry point rejects codes outside the enum with INVALID_VALUE
    assert parament.paramentlib.lib.Parament_setPrecisionMode(None, 1) == 50


# ---- additions: argument checks of the ensemble and gradient entry points (host only) ----

def _raw(name, restype=ctypes.c_int):
    """The entry point without the ctypes argument conversions, so that null pointers and
    out-of-range values reach the C checks."""
    from parament_amd import paramentlib
    f = getattr(ctypes.CDLL(paramentlib.lib._name), name)
    f.restype = restype
    return f


def test_ensemble_create_rejects_bad_arguments():
    create = _raw("Parament_ensembleCreate")
    u = lambda *v: (ctypes.c_uint * len(v))(*v)
    dbl = lambda *v: (ctypes.c_double * len(v))(*v)
    none, midpoint, simpson = 0, 0x01000000, 0x02000000
    dims, amps, pts, dts = u(4, 8), u(1, 2), u(9, 9), dbl(0.01, 0.02)
    cases = {
        "empty ensemble": (0, dims, amps, pts, dts, none),
        "null dims": (2, None, amps, pts, dts, none),
        "null dt": (2, dims, amps, pts, None, none),
        "unknown quadrature": (2, dims, amps, pts, dts, 7),
        "Magnus flag": (2, dims, amps, pts, dts, simpson | 0x10000000),
        "dim 0": (2, u(4, 0), amps, pts, dts, none),
        "no controls": (2, dims, u(1, 0), pts, dts, none),
        "more controls than a slice holds": (2, dims, u(1, 257), pts, dts, none),
        "dt 0": (2, dims, amps, pts, dbl(0.01, 0.0), none),
        "dt < 0": (2, dims, amps, pts, dbl(-0.01, 0.02), none),
        "dt NaN": (2, dims, amps, pts, dbl(0.01, float("nan")), none),
        "no samples": (2, dims, amps, u(9, 0), dts, none),
        "midpoint of one sample": (2, dims, amps, u(9, 1), dts, midpoint),
        "Simpson of two samples": (2, dims, amps, u(9, 2), dts, simpson),
    }
    for what, args in cases.items():
        h = ctypes.c_void_p()
        assert create(ctypes.byref(h), *args) == 50, what
        assert not h.value, what
    assert create(None, 2, dims, amps, pts, dts, none) == 50


def test_ensemble_entry_points_reject_null_handle():
    assert _raw("Parament_ensembleDestroy")(None) == 0
    for name, nargs in (("Parament_ensembleSetHamiltonians", 1), ("Parament_ensembleSetExecution", 1),
                        ("Parament_ensembleTune", 0), ("Parament_ensembleGetTuning", 1),
                        ("Parament_ensembleEquiprop", 2), ("Parament_ensembleEquipropDevice", 2),
                        ("Parament_ensembleSynchronize", 0), ("Parament_ensembleGateFidelityGradient", 4),
                        ("Parament_ensembleGateFidelityGradientDevice", 4)):
        assert _raw(name)(None, *([None] * nargs)) == 50, name


@pytest.mark.parametrize("sfx,cplx", [("", np.complex64), ("_fp64", np.complex128)])
def test_gradient_entry_points_resolve_on_host(sfx, cplx):
    """No context, or a context without a Hamiltonian: the gradient returns the reference's
    codes (INVALID_VALUE 50, NO_HAMILTONIAN 80) before any device work, and the workspace query
    reads zero."""
    grad = _raw("Parament_gateFidelityGradient" + sfx)
    grad.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_uint, ctypes.c_uint,
                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    real = np.float32 if cplx == np.complex64 else np.float64
    c = np.zeros(8, cplx)
    t = np.eye(2, dtype=cplx)
    f = np.zeros(1, real)
    g = np.zeros(8, real)
    ptr = lambda a: a.ctypes.data
    assert grad(None, ptr(c), 0.1, 8, 1, ptr(t), ptr(f), ptr(g)) == 50
    create, destroy = _raw("Parament_create" + sfx), _raw("Parament_destroy" + sfx)
    h = ctypes.c_void_p()
    assert create(ctypes.byref(h)) == 0
    try:
        assert grad(h, ptr(c), 0.1, 8, 1, ptr(t), ptr(f), ptr(g)) == 80
        workspace = _raw("Parament_getGradientWorkspace" + sfx, ctypes.c_size_t)
        workspace.argtypes = [ctypes.c_void_p]
        assert workspace(h) == 0 and workspace(None) == 0
    finally:
        assert destroy(h) == 0


def test_ensemble_class_checks_arguments():
    from parament_amd.ensemble import Ensemble
    with pytest.raises(ValueError, match="same length"):
        Ensemble([4, 8], [1], [9, 9], 0.01)
    with pytest.raises(ValueError, match="Unknown quadrature"):
        Ensemble([4], [1], [9], 0.01, "magnus")
    with pytest.raises(ValueError, match="Invalid value"):
        Ensemble([4, 0], [1, 1], [9, 9], 0.01)
    with pytest.raises(ValueError, match="Invalid value"):
        Ensemble([4], [1], [2], 0.01, "simpson")


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and not r.stdout.strip()


def test_coefficient_limit_resolves_on_host():
    """At most 256 operator coefficients per slice: A controls, with Magnus A + A + A(A-1)/2
    (20 controls fit, 22 do not); checked before any device work."""
    import parament_amd as parament
    eye = np.eye(2)
    with parament.Parament() as p:
        with pytest.raises(ValueError, match="Invalid value"):
            p.set_hamiltonian(eye, *([eye] * 257))
        with pytest.raises(ValueError, match="Invalid value"):
            p.set_hamiltonian(eye, *([eye] * 22), use_magnus=True, quadrature_mode="simpson")
