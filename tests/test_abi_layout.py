"""The ctypes structures of the Python binding have the size and field offsets
of the C structures they mirror (include/metapart_amd.h and the generator's
mp_synth.h), as gcc lays them out: a field added on one side only fails here,
not as garbage read on the device."""
import ctypes as ct
import os
import shutil
import subprocess

import pytest

import metapart_amd as mp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PAIRS = [(mp.Cfg, "mp_cfg"), (mp.ScanStats, "mp_scan_stats"), (mp.FinishStats, "mp_finish_stats"),
         (mp.SampleResult, "mp_sample_result"), (mp.ServerCfg, "mp_server_cfg"), (mp.JobResult, "mp_job_result"),
         (mp.SynthCfg, "mp_synth_cfg")]


def _c_layout(tmp):
    lines = ["#include <stddef.h>", "#include <stdio.h>", "#include <stdint.h>", '#include "metapart_amd.h"',
             '#include "mp_synth.h"', "int main(void) {"]
    for cls, cname in PAIRS:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in cls._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ["return 0;", "}"]
    src = os.path.join(tmp, "layout.c")
    open(src, "w").write("\n".join(lines) + "\n")
    exe = os.path.join(tmp, "layout")
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "metapart_amd", "csrc"), src, "-o", exe], check=True)
    out = {}
    for ln in subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout.splitlines():
        a, b, c = ln.split()
        out[(a, b)] = int(c)
    return out


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structures_match_the_c_layout(tmp_path):
    lay = _c_layout(str(tmp_path))
    for cls, cname in PAIRS:
        assert ct.sizeof(cls) == lay[(cname, "size")], cname
        for f, _ in cls._fields_:
            assert getattr(cls, f).offset == lay[(cname, f)], (cname, f)
