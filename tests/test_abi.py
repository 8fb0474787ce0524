"""The C-ABI library loads on a machine without a GPU and exports every entry point
include/kc.h declares; the ctypes binding covers exactly that set."""
import ctypes
import os
import re

from conftest import ROOT


def header_functions():
    with open(os.path.join(ROOT, "include", "kc.h")) as fh:
        text = fh.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    from kc import _abi
    lib = _abi.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_abi.EXPORTS) == names


def test_struct_layouts_match_header():
    from kc import _abi
    # kc_opts: int32 device, uint32 k, uint64 bloom_bits, double, int32, float, uint32
    assert ctypes.sizeof(_abi.kc_opts) == 40
    assert _abi.kc_opts.bloom_bits.offset == 8 and _abi.kc_opts.min_coverage.offset == 16
    assert ctypes.sizeof(_abi.kc_info) == 8 * 8 + 4 * 2 + 8 * 2 + 8
    assert ctypes.sizeof(_abi.kc_tune) == 40 and _abi.kc_tune.ms.offset == 32
    assert ctypes.sizeof(_abi.kc_packed_info) == 64
    assert ctypes.sizeof(_abi.kc_unitigs_info) == 64 and _abi.kc_unitigs_info.ms_stage.offset == 32


def test_struct_layouts_match_c_compiler(tmp_path):
    """The ctypes layouts equal what a C compiler makes of include/kc.h."""
    import subprocess
    from kc import _abi
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "kc.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(kc_opts), sizeof(kc_info),'
                   ' sizeof(kc_tune), offsetof(kc_opts, table_load), offsetof(kc_info, cutoff),'
                   ' offsetof(kc_info, ms_stage2), offsetof(kc_tune, table_load), sizeof(kc_packed_info),'
                   ' offsetof(kc_packed_info, ms_device)); return 0;}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [ctypes.sizeof(_abi.kc_opts), ctypes.sizeof(_abi.kc_info), ctypes.sizeof(_abi.kc_tune),
                   _abi.kc_opts.table_load.offset, _abi.kc_info.cutoff.offset, _abi.kc_info.ms_stage2.offset,
                   _abi.kc_tune.table_load.offset, ctypes.sizeof(_abi.kc_packed_info),
                   _abi.kc_packed_info.ms_device.offset]


def test_no_device_is_an_error_not_a_crash():
    """Without a GPU the constructors return a status (no fallback path exists)."""
    import torch
    from kc import _abi
    if torch.cuda.is_available():
        return
    h = ctypes.c_void_p()
    o = _abi.kc_opts(device=0, k=21, bloom_bits=1 << 20, min_coverage=5.0)
    assert _abi.lib().kc_create(ctypes.byref(h), ctypes.byref(o)) != 0
    assert _abi.lib().kc_filter_create(ctypes.byref(h), 0, 1 << 20) != 0
    assert _abi.lib().kc_create(ctypes.byref(h), None) == _abi.KC_EINVAL
    # k is checked before the device: 1..127 supported, 128 valid for Kmer256 but not here
    for k, want in ((0, _abi.KC_EINVAL), (128, _abi.KC_EUNSUPPORTED), (129, _abi.KC_EINVAL), (127, _abi.KC_EHIP)):
        o = _abi.kc_opts(device=0, k=k, bloom_bits=1 << 20, min_coverage=5.0)
        assert _abi.lib().kc_create(ctypes.byref(h), ctypes.byref(o)) == want, k
