"""The C-ABI library loads without a GPU and exports what include/volrus.h declares."""
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(vr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_symbols():
    assert len(declared_symbols()) >= 20


def test_library_exports_every_declared_symbol(vr):
    lib = vr.load_library()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_python_binding_covers_header(vr):
    assert set(declared_symbols()) <= set(vr._SIGS), sorted(set(declared_symbols()) - set(vr._SIGS))


def test_rust_binding_declares_every_entry_point():
    """The reference-side (Rust) extern block in the integration notes names
    every function of the header, with its own parameter count."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    header = open(os.path.join(ROOT, "include", "volrus.h")).read()
    for name in declared_symbols():
        m = re.search(r"pub fn " + name + r"\(([^)]*)\)", text)
        assert m, f"{name} has no Rust declaration"
        c = re.search(r"^[A-Za-z_][\w *]*\b" + name + r"\(([^)]*)\)", header, re.M)  # the declaration
        n_rust = len([a for a in m.group(1).split(",") if a.strip()])
        n_c = len([a for a in c.group(1).split(",") if a.strip() and a.strip() != "void"])
        assert n_rust == n_c, (name, n_rust, n_c)


def test_no_cpu_fallback_in_product():
    """The product package never reaches the oracle."""
    pkg = os.path.join(ROOT, "sharc.volrus_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".hpp")):
                text = open(os.path.join(dirpath, f), errors="ignore").read()
                assert "oracle" not in text.lower(), f"{f} mentions the oracle"


# ---------------------------------------------------------------- host stubs vs device code
def _elf_sections(data):
    """(name, offset, size, type, link) of each section of a 64-bit little-endian ELF."""
    import struct

    shoff, = struct.unpack_from("<Q", data, 0x28)
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", data, 0x3A)
    raw = [struct.unpack_from("<IIQQQQIIQQ", data, shoff + i * shentsize) for i in range(shnum)]
    stro = raw[shstrndx][4]

    def name(off):
        end = data.index(b"\0", stro + off)
        return data[stro + off:end].decode()

    return [(name(r[0]), r[4], r[5], r[1], r[6]) for r in raw]


def _elf_symbols(data):
    import struct

    secs = _elf_sections(data)
    out = set()
    for nm, off, size, typ, link in secs:
        if typ not in (2, 11):  # SYMTAB, DYNSYM
            continue
        stroff = secs[link][1]
        for i in range(size // 24):
            st_name, = struct.unpack_from("<I", data, off + i * 24)
            end = data.index(b"\0", stroff + st_name)
            out.add(data[stroff + st_name:end].decode(errors="replace"))
    return out


def _device_kernels(lib_bytes):
    """Kernel names of every gfx950 code object in the library's offload bundles."""
    import struct

    sec = [s for s in _elf_sections(lib_bytes) if s[0] == ".hip_fatbin"]
    assert sec, "no .hip_fatbin section"
    _, off, size, _, _ = sec[0]
    fat = lib_bytes[off:off + size]
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    kernels, pos = set(), fat.find(magic)
    while pos >= 0:
        n, = struct.unpack_from("<Q", fat, pos + 24)
        q = pos + 32
        for _ in range(n):
            eoff, esize, tlen = struct.unpack_from("<QQQ", fat, q)
            triple = fat[q + 24:q + 24 + tlen].decode()
            q += 24 + tlen
            if "amdgcn" in triple:
                assert "gfx950" in triple, triple
                co = fat[pos + eoff:pos + eoff + esize]
                kernels |= {s[:-3] for s in _elf_symbols(co) if s.endswith(".kd")}
        pos = fat.find(magic, pos + 1)
    return kernels


def test_every_kernel_stub_has_device_code():
    """Each kernel the host side can launch exists, under the same mangled
    name, in the gfx950 code object (a parameter type that differs between
    the host and device passes would leave the launch without a kernel)."""
    data = open(os.path.join(ROOT, "sharc.volrus_amd", "libvolrus_amd.so"), "rb").read()
    device = _device_kernels(data)
    assert len(device) > 20
    stub = "__device_stub__"
    host = set()
    for s in _elf_symbols(data):
        m = re.match(r"_Z(\d+)" + stub, s)
        if m:
            ident_len = int(m.group(1))
            base = s[2 + len(m.group(1)) + len(stub):2 + len(m.group(1)) + ident_len]
            rest = s[2 + len(m.group(1)) + ident_len:]
            host.add(f"_Z{len(base)}{base}{rest}")
    assert len(host) > 20
    missing = sorted(host - device)
    assert not missing, f"kernels launched from the host without device code: {missing[:5]}"
