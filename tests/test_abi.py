"""The C ABI (CPU only): librtg.so loads without a GPU, exports every function declared in
include/*.h, and reports errors through status codes and rtg_last_error."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

from conftest import ROOT, load_rtg

rtg = load_rtg()


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if "inline" in m.group(0) or "template" in m.group(0):
                continue  # C++ convenience wrappers defined in the header
            if name not in {"if", "while", "for", "return", "sizeof"} and not name.startswith("RTG_"):
                names.add(name)
    return names


def test_header_symbols_exported():
    lib = ctypes.CDLL(str(rtg.lib_path()))
    names = declared_functions()
    assert "rtg_tc_count" in names and "rtg_graph_from_pairs" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_error_reporting_without_gpu_or_bad_args():
    lib = ctypes.CDLL(str(rtg.lib_path()))
    lib.rtg_last_error.restype = ctypes.c_char_p
    h = ctypes.c_void_p()
    rc = lib.rtg_create(-1, ctypes.byref(h))
    assert rc != 0 and lib.rtg_last_error()
    assert lib.rtg_gen_ba(10, 0, ctypes.c_uint64(1), None, None) == 1  # RTG_EINVAL
    assert b"n > m" in lib.rtg_last_error()


def test_struct_sizes_match_header():
    # rtg_tc_stats layout: 2 doubles, 2 u64, i64, i32, 8 i32, 8 i64, 8 doubles, 2 i32, u64
    assert ctypes.sizeof(rtg.TcStats) == 8 * 2 + 8 * 2 + 8 + 4 + 4 * 8 + 4 + 8 * 8 + 8 * 8 + 4 * 2 + 8
    assert ctypes.sizeof(rtg.TcOpts) == 4 + 4 + 8 + 8 + 8 + 4 + 4


def test_compat_header_templates_compile_and_link(tmp_path):
    """A reference-style driver that calls every drop-in of include/rtg_compat.h builds and links
    against librtg (the set-intersection baselines are header templates over the reference's
    Dataset field names); it is compiled, not run."""
    import shutil
    import subprocess
    cxx = shutil.which("g++") or shutil.which("clang++")
    if cxx is None:
        import pytest
        pytest.skip("no host C++ compiler")
    src = tmp_path / "driver.cpp"
    src.write_text(r'''
#include <vector>
#include "rtg_compat.h"
struct Dataset {  // set_intersection/utils/util.h:19-28 field names
    unsigned num_of_sets_a, num_of_sets_b, num_of_sets, max_element, intersection;
    std::vector<unsigned> set_offsets, elements;
};
int main(int argc, char **) {
    Dataset d{1, 1, 2, 10, 1, {0, 2, 4}, {1, 5, 5, 7}};
    std::vector<unsigned> r;
    if (argc > 99) {  // never taken: instantiate and link every drop-in without running it
        BinarySearchBasedIntersection(d, 1, 256, r);
        ThreadBinarySearch(d, 1, 256, r);
        HashBasedIntersection(d, 1, 256, 1024, 16, r);
        IntersectPathBasedIntersection(d, 1, 256, r);
        BitmapNaiveBasedIntersection(d, 1, 256, r);
        BitmapDynamicBasedIntersection(d, 1, 256, r);
        int adjs[1] = {0}, offs[2] = {0, 0}, data[1] = {1}, q[1] = {1}, out[1];
        CountOnGPU(adjs, offs, 1, 0, 0);
        BSAccessCountFunction(data, 1, q, 1, out);
    }
    return 0;
}
''')
    lib = rtg.lib_path()
    exe = tmp_path / "driver"
    proc = subprocess.run([cxx, "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                           f"-L{lib.parent}", "-lrtg", f"-Wl,-rpath,{lib.parent}"], capture_output=True, text=True)
    assert proc.returncode == 0, proc.stderr
    assert subprocess.run([str(exe)]).returncode == 0
