"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/sfeng.h declares, and rejects bad arguments with a status code (no
GPU work is started by these calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "sfeng.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sf_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from segfault_py_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.fail("libsfeng.so is not built; run python -m segfault_py_amd.build")
    return engine.lib()


def test_exports_every_declared_symbol(lib):
    from segfault_py_amd import engine
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(engine.exported_symbols()) == names


def test_version_and_error_string(lib):
    assert lib.sf_version() == 200
    assert isinstance(lib.sf_last_error(), bytes)


def test_argument_errors_are_status_codes(lib):
    from segfault_py_amd import engine
    i = np.zeros(4, dtype=np.int32)
    d = np.zeros(4)
    # no functions
    assert lib.sf_eri_components(0, i, i, d, d, d, np.zeros(12, dtype=np.int32), np.zeros(12), 0,
                                 i, d) == 1
    assert b"bad arguments" in lib.sf_last_error()
    # total angular momentum beyond 7 Rys nodes (L = 16 > 13)
    shells = np.array([4, 0, 0] * 4, dtype=np.int32)
    q = np.arange(4, dtype=np.int32)
    rc = lib.sf_eri_components(4, np.ones(4, dtype=np.int32), np.arange(4, dtype=np.int32),
                               np.ones(4), np.ones(4), np.ones(4), shells, np.zeros(12), 1, q,
                               np.zeros(1))
    assert rc == 5
    # g shells are outside the J/K engine (s, p, d, f)
    h = ctypes.c_void_p()
    rc = lib.sf_basis_create(1, np.array([4], dtype=np.int32), np.array([1], dtype=np.int32),
                             np.zeros(3), np.ones(1), np.ones(1), 1, np.zeros(1, dtype=np.int32),
                             ctypes.byref(h))
    assert rc == 5 and not h.value
    with pytest.raises(engine.EngineError):
        engine.check(rc)


def test_shell_atom_must_be_in_its_molecule(lib):
    """A shell's atom index outside its own molecule's nuclei is rejected
    (status 1, no handle) before any device work: the gradient addresses
    per-atom rows and lists through it."""
    h = ctypes.c_void_p()
    L = np.zeros(2, dtype=np.int32)
    npr = np.ones(2, dtype=np.int32)
    for atoms in ([0, 2], [-1, 0], [0, 5]):
        rc = lib.sf_basis_create(2, L, npr, np.zeros(6), np.ones(2), np.ones(2), 2,
                                 np.array(atoms, dtype=np.int32), ctypes.byref(h))
        assert rc == 1 and not h.value, atoms
        assert b"outside molecule" in lib.sf_last_error()
    # batch: molecule 1 owns nuclei [1, 2); its shell may not name nucleus 0
    rc = lib.sf_basis_create_batch(2, np.array([0, 1, 2], dtype=np.int32), 2, L, npr, np.zeros(6),
                                   np.ones(2), np.ones(2), 2, np.array([0, 1, 2], dtype=np.int32),
                                   np.array([0, 0], dtype=np.int32), ctypes.byref(h))
    assert rc == 1 and not h.value


def test_wrappers_check_tensors_before_the_library(monkeypatch):
    """The Python wrappers reject CPU, float32, wrongly shaped and (for
    outputs) non-contiguous tensors with EngineError before any library
    call: a bad pointer reaching a kernel would fault the GPU."""
    import torch
    from segfault_py_amd import engine

    class NoLib:
        def __getattr__(self, name):
            raise AssertionError(f"library called: {name}")

    monkeypatch.setattr(engine, "_torch", lambda: torch)
    monkeypatch.setattr(engine, "lib", lambda: NoLib())
    b = engine.Basis.__new__(engine.Basis)
    b.shape = (4, 4)
    b._h = None
    b.mol = type("M", (), {"natom": 1})()
    good = torch.zeros(4, 4, dtype=torch.float64)   # CPU: still rejected (not CUDA)
    for bad in (good, good.float(), torch.zeros(3, 3, dtype=torch.float64), None):
        with pytest.raises(engine.EngineError):
            b.jk(bad)
        with pytest.raises(engine.EngineError):
            b.gradient(good, bad)
        with pytest.raises(engine.EngineError):
            b._check_out((bad, bad))
    with pytest.raises(engine.EngineError):
        b._check_out(good)
    with pytest.raises(engine.EngineError):
        engine.boys(torch.zeros(4, dtype=torch.float32), 2)
    with pytest.raises(engine.EngineError):
        engine.rys_roots(torch.zeros(4, dtype=torch.float64), 2)   # CPU tensor


def test_batch_and_graph_argument_errors(lib):
    """Bad molecule ranges, tuning shapes and null handles are rejected
    before any device work."""
    from segfault_py_amd import engine
    h = ctypes.c_void_p()
    L = np.zeros(2, dtype=np.int32)
    npr = np.ones(2, dtype=np.int32)
    # molecule shell ranges must cover [0, nshell] with no empty molecule
    for ms in ([0, 2, 2], [1, 2], [0, 1]):
        ms = np.array(ms, dtype=np.int32)
        rc = lib.sf_basis_create_batch(len(ms) - 1, ms, 2, L, npr, np.zeros(6), np.ones(2),
                                       np.ones(2), 2, np.array([0, 1, 2][:len(ms)], dtype=np.int32),
                                       np.zeros(2, dtype=np.int32), ctypes.byref(h))
        assert rc == 1 and not h.value, ms
    # nuclei ranges beyond natom
    rc = lib.sf_basis_create_batch(2, np.array([0, 1, 2], dtype=np.int32), 2, L, npr, np.zeros(6),
                                   np.ones(2), np.ones(2), 1, np.array([0, 1, 2], dtype=np.int32),
                                   np.zeros(2, dtype=np.int32), ctypes.byref(h))
    assert rc == 1 and not h.value
    assert lib.sf_basis_create_batch(0, np.zeros(1, dtype=np.int32), 2, L, npr, np.zeros(6),
                                     np.ones(2), np.ones(2), 0, np.zeros(1, dtype=np.int32),
                                     np.zeros(2, dtype=np.int32), ctypes.byref(h)) == 1
    shapes = np.full(9 * engine.NUM_QUARTET_CLASSES, 256, dtype=np.int32)
    assert lib.sf_set_tuning(None, shapes) == 1
    assert lib.sf_set_streams(None, 1) == 1
    assert lib.sf_jk_graph(None, None, None, None, 0, None) == 1
    assert lib.sf_autotune(None, None, 0, None, shapes, np.zeros(3 * engine.NUM_QUARTET_CLASSES)) == 1
    assert lib.sf_basis_set_layout(None, np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int32)) == 1


def test_product_never_touches_the_oracle():
    """The package, the C-ABI header and bench's timed legs do not import,
    link or load anything from oracle/ (only tests, smoke() and bench's
    cpu_baseline leg use it, as the checker)."""
    pkg = os.path.join(ROOT, "segfault_py_amd")
    pat = re.compile(r"\boracle\b|liboracle", re.I)
    hits = []
    for d, _, files in os.walk(pkg):
        if "_build" in d or "__pycache__" in d:
            continue
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".h", ".cpp", ".inc")):
                p = os.path.join(d, f)
                for n, line in enumerate(open(p, errors="replace"), 1):
                    if pat.search(line):
                        hits.append(f"{p}:{n}: {line.strip()}")
    hits += [f"include/sfeng.h: {l.strip()}" for l in open(os.path.join(ROOT, "include", "sfeng.h"))
             if "liboracle" in l or "#include" in l and "oracle" in l]
    assert not hits, "\n".join(hits)
    # bench.py imports the oracle only inside its cpu_baseline function
    src = open(os.path.join(ROOT, "bench.py")).read()
    funcs = re.split(r"\ndef ", src)
    users = [f.split("(")[0] for f in funcs[1:] if "import oracle" in f]
    assert users == ["cpu_baseline"], users
    assert "import oracle" not in funcs[0]


def test_no_cpu_fallback(monkeypatch):
    """Without the library, or without a GPU, the product path raises
    EngineError instead of computing anything on the CPU."""
    import torch
    from segfault_py_amd import engine
    monkeypatch.setattr(engine, "_LIB", None)
    monkeypatch.setattr(engine, "LIB_PATH", os.path.join(ROOT, "no_such_dir", "libsfeng.so"))
    with pytest.raises(engine.EngineError, match="not built"):
        engine.lib()
    monkeypatch.undo()
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: the no-GPU error path cannot be exercised here")
    with pytest.raises(engine.EngineError, match="no GPU"):
        engine.boys(torch.zeros(4, dtype=torch.float64), 2)
