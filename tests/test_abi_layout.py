"""The ctypes mirrors of the ABI structs have the C layout: every struct's
size and every field's offset, from a C program compiled against
include/menos_hip.h on the CPU."""
import ctypes
import json
import os
import subprocess

import pytest

from menos_amd import hip as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MIRRORS = {
    "mh_device_info": H.DeviceInfo, "mh_circuit": H.CCircuit, "mh_execution_request": H.CRequest,
    "mh_execution_result": H.CResult, "mh_cost_estimate": H.CCost, "mh_calibration": H.Calibration,
    "mh_batch_request": H.CBatchRequest, "mh_batch_result": H.CBatchResult,
    "mh_server_options": H.CServerOptions, "mh_job": H.CJob, "mh_job_result": H.CJobResult,
    "mh_server_stats": H.CServerStats, "mh_tier_pass": H.CTierPass, "mh_tier_request": H.CTierRequest,
    "mh_tier_cost": H.CTierCost,
}


def _c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "menos_hip.h"', 'int main(void) {',
             '  printf("{");']
    first = True
    for cname, mirror in MIRRORS.items():
        sep = "" if first else ","
        first = False
        lines.append(f'  printf("{sep}\\"{cname}\\": [%zu", sizeof({cname}));')
        for field, _ in mirror._fields_:
            lines.append(f'  printf(", %zu", offsetof({cname}, {field}));')
        lines.append('  printf("]");')
    lines += ['  printf("}\\n");', '  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)], check=True)
    return json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


def test_ctypes_mirrors_match_the_c_layout(tmp_path):
    c = _c_layout(tmp_path)
    for cname, mirror in MIRRORS.items():
        got = [ctypes.sizeof(mirror)] + [getattr(mirror, f).offset for f, _ in mirror._fields_]
        assert got == c[cname], cname
