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
    n_decl = re.search(r"let sources: \[\(&str, &\[&str\]\); (\d+)\]", text)
    assert n_decl and int(n_decl.group(1)) == len(pairs), "array length must match its entries"
    m = re.search(r"let uniform_skip: &\[&str\] = &\[(.*?)\];", text)
    assert m and re.findall(r'"([^"]+)"', m.group(1)) == build.UNIFORM_SKIP
    flags = re.findall(r'"([^"]+)"', re.search(r"let flags = \[(.*?)\];", text, re.S).group(1))
    assert flags == [f for f in build.FLAGS if not f.startswith("-Wno-")]


def _layout(fields):
    """[(name, size, count)] -> {name: offset} under C / repr(C) rules
    (each field aligned to its element size)."""
    off, out = 0, {}
    for name, size, count in fields:
        off = (off + size - 1) // size * size
        out[name] = off
        off += size * count
    align = max(size for _, size, _ in fields)
    return out, (off + align - 1) // align * align


def _rust_fields():
    text = open(os.path.join(ROOT, "fast-atari-rs_amd", "rust", "hip_env.rs")).read()
    body = re.search(r"pub struct AtariStateGpu \{(.*?)\n\}", text, re.S).group(1)
    fields = []
    for name, ty in re.findall(r"pub (\w+): ([^,\n]+)", body):
        ty = ty.strip()
        m = re.fullmatch(r"\[(u8|u16); (\d+)\]", ty)
        if m:
            fields.append((name, 1 if m.group(1) == "u8" else 2, int(m.group(2))))
        else:
            fields.append((name, {"u8": 1, "u16": 2}[ty], 1))
    return fields


def _c_fields():
    text = open(HEADER).read()
    body = re.search(r"typedef struct AtariStateGpu \{(.*?)\} AtariStateGpu;", text, re.S).group(1)
    fields = []
    for ty, names in re.findall(r"(uint8_t|uint16_t)\s+([^;]+);", body):
        for nm in names.split(","):
            nm = nm.strip()
            m = re.fullmatch(r"(\w+)\[(\d+)\]", nm)
            size = 1 if ty == "uint8_t" else 2
            fields.append((m.group(1), size, int(m.group(2))) if m else (nm, size, 1))
    return fields


def test_rust_state_struct_matches_the_c_header_and_the_reference_names():
    """hip_env.rs's AtariStateGpu has the C header's fields in the same order,
    with the same names, sizes and offsets (the reference's struct,
    src/cuda_env.rs:12-38), and 222 bytes in all; the Python dtype agrees."""
    import fastatari
    rust, c = _rust_fields(), _c_fields()
    assert [f[0] for f in rust] == [f[0] for f in c]
    assert rust == c
    lr, size_r = _layout(rust)
    lc, size_c = _layout(c)
    assert lr == lc and size_r == size_c == 222
    for name in fastatari.STATE_DTYPE.names:
        if name in lc:
            assert fastatari.STATE_DTYPE.fields[name][1] == lc[name], name
    for name in ("cxm0p", "cxppmm", "pos_p0", "pos_bl", "hmp0", "hmbl", "_pad1", "_pad3"):
        assert name in lr
    assert (lr["cxm0p"], lr["pos_p0"], lr["hmbl"], lr["bank"]) == (160, 176, 190, 217)


def test_rust_binding_checks_lengths_before_each_ffi_call():
    """Every slice handed to the library is checked against the env count
    first (the reference asserts the same lengths, src/cuda_env.rs:145,167)."""
    text = open(os.path.join(ROOT, "fast-atari-rs_amd", "rust", "hip_env.rs")).read()
    for fn in ("pub fn step(&mut self, actions: &[u8])", "pub fn step_multi", "pub fn step_frameskip",
               "pub fn upload_states", "pub fn new(roms: &[Vec<u8>], counts: &[usize])", "pub fn tune"):
        i = text.index(fn)
        body = text[i: text.index("\n    }\n", i)]
        assert "assert_eq!" in body, fn
    assert "pub fn from_headless(atari: &HeadlessAtari) -> Self" in text
