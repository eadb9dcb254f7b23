"""The C-ABI library loads without a GPU and exports every function that
include/fastatari.h declares, with a ctypes signature for each."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fastatari.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_reference_surface():
    names = declared()
    for n in ("fa_create", "fa_destroy", "fa_set_sorted", "fa_reset", "fa_step", "fa_step_multi",
              "fa_download_states", "fa_num_envs", "fa_last_error"):
        assert n in names
    assert len(names) >= 50


def test_library_exports_every_declared_symbol():
    import fastatari
    lib = ctypes.CDLL(fastatari.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_every_symbol_has_a_binding():
    import fastatari
    L = fastatari.load_library()
    for n in declared():
        fn = getattr(L, n)
        assert fn.argtypes is not None, f"{n} has no ctypes signature"


def test_null_handles_are_errors_not_crashes():
    import fastatari
    L = fastatari.load_library()
    assert L.fa_set_sorted(None, 1) == -1
    assert b"NULL" in L.fa_last_error()
    assert L.fa_num_envs(None) == 0
    assert L.fa_batch_num_envs(None) == 0
    assert L.fa_server_poll(None, 0) == -1
    assert L.fa_vecenv_step(None, 0, None, None, None, None) == -1
    assert L.fa_vecenv_num_envs(None) == 0
    assert L.fa_vecenv_set_frameskip(None, 4) == -1
    L.fa_destroy(None)
    L.fa_batch_destroy(None)
    L.fa_server_destroy(None)
    L.fa_vecenv_destroy(None)


def test_rust_build_script_compiles_what_the_python_build_compiles():
    """rust/build.rs (no Rust toolchain here) names the same translation units
    with the same extra options as fastatari/build.py."""
    from fastatari import build
    text = open(os.path.join(ROOT, "fast-atari-rs_amd", "rust", "build.rs")).read()
    pairs = re.findall(r'\("(fa_[a-z_]+\.hip)",\s*(uniform_skip|&\[\])\)', text)
    rust = {src: (build.UNIFORM_SKIP if opt == "uniform_skip" else []) for src, opt in pairs}
    assert rust == {src: list(extra) for src, extra in build.SOURCES}
    m = re.search(r"let uniform_skip: &\[&str\] = &\[(.*?)\];", text)
    assert m and re.findall(r'"([^"]+)"', m.group(1)) == build.UNIFORM_SKIP
    flags = re.findall(r'"([^"]+)"', re.search(r"let flags = \[(.*?)\];", text, re.S).group(1))
    assert flags == [f for f in build.FLAGS if not f.startswith("-Wno-")]
